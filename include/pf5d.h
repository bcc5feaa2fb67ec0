/* pf5d.h -- C ABI of the MI355X (gfx950) phasefield5d explicit-step library.
 *
 * All compute entry points take DEVICE pointers and a hipStream_t (passed as
 * void*; NULL = the default stream) and return 0 on success or a negative
 * PF5D_E* code; pf5d_last_error() gives the message of the last failure on the
 * calling thread.  Arrays use the reference layout: C-order, the four
 * independent components (Mn, Ni, Co, Cu) innermost.  `prec` selects the
 * element type of device arrays: PF5D_F64 (double) or PF5D_F32 (float).
 *
 * Reference interfaces each entry point replaces are cited per declaration
 * (paths relative to the phasefield5d repository).
 */
#ifndef PF5D_H
#define PF5D_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PF5D_F64 0
#define PF5D_F32 1

#define PF5D_OK 0
#define PF5D_EINVAL (-1)   /* bad argument (maps to ValueError) */
#define PF5D_EDEVICE (-2)  /* HIP runtime error (RuntimeError) */
#define PF5D_ENOTIMPL (-3) /* unsupported dimension (NotImplementedError) */
#define PF5D_ENOMEM (-4)   /* device allocation failed (MemoryError) */

typedef struct pf5d_table pf5d_table;
typedef struct pf5d_elastic pf5d_elastic;
typedef struct pf5d_engine pf5d_engine;

const char* pf5d_last_error(void);
int pf5d_version(void);
/* device properties used for launch sizing: [CU count, max LDS per block, warp size] */
int pf5d_device_info(int out[3]);

/* ---- CALPHAD table (solver/interpolation.py:13-35 build_4d_grid output) ----
 * grid: host (num,num,num,num,nprops) float64 with NaN outside the simplex,
 * nprops == 9.  The library keeps a padded device copy in `prec`, and a cube
 * table of the linear region (per hypercube its 16 corners contiguous, the
 * order the interpolation reads them) for gathers that miss the LDS node box. */
int pf5d_table_create(int prec, int num, int nprops, const double* grid_host, double resolution,
                      pf5d_table** out);
int pf5d_table_destroy(pf5d_table* t);

/* ---- per-function drop-in kernels (one reference function each) ---- */

/* interpolator_nb -- solver/interpolation.py:154-211.  X (ncell,4) -> out rows of
 * stride ostride (>= 9): linear where sum(X) <= 1-4h-1e-9, nearest node else. */
int pf5d_interpolate(int prec, const void* X, long ncell, const pf5d_table* tab, void* out, long ostride,
                     void* stream);

/* calculate_laplacian -- solver/operators.py:10-81.  shape[dim], periodic. */
int pf5d_laplacian(int prec, const void* X, int dim, const long* shape, double cell_size, void* out,
                   void* stream);

/* composition_laplacian *= kappa; mu -= composition_laplacian -- examples/simulate.py:254-255.
 * mu rows of stride mstride (the interpolated-data view). */
int pf5d_kappa_subtract(int prec, void* lap, void* mu, long mstride, long ncell, const double kappa[4],
                        void* stream);

/* get_divergence_from_mu_function(dim)(X, Mob, mu, inv_dx, out) -- solver/fluxes.py:296-416.
 * mob / mu: rows of stride bstride / mstride (views into the interpolated data). */
int pf5d_divergence_from_mu(int prec, const void* X, const void* mob, long bstride, const void* mu, long mstride,
                            int dim, const long* shape, double inv_dx, void* out, void* stream);

/* calculate_gradients_pm(array, cell_size, grad_p, grad_m) -- solver/operators.py:88-158.
 * A: rows of 4 components with row stride astride (e.g. the mu view of the interpolated data);
 * grad_p / grad_m: (dim, ncell, 4), forward (A(c+e_d) - A(c))/h and backward (A(c) - A(c-e_d))/h. */
int pf5d_gradients_pm(int prec, const void* A, long astride, int dim, const long* shape, double h, void* grad_p,
                      void* grad_m, void* stream);

/* get_flux_function(dim)(X, Mob, grad_mu_p, grad_mu_m, fluxes_p, fluxes_m) -- solver/fluxes.py:6-125.
 * mob rows of stride bstride; gradients and fluxes (dim, ncell, 4).  Used for the save-step
 * total_flux diagnostic (examples/simulate.py:302-312, 323). */
int pf5d_fluxes_pm(int prec, const void* X, const void* mob, long bstride, int dim, const long* shape,
                   const void* grad_p, const void* grad_m, void* fluxes_p, void* fluxes_m, void* stream);

/* Save-step field statistics of an (ncell, ncomp <= 8) array with row stride astride, in FP64:
 * out[0..n) sums (compensated), [n..2n) minima, [2n..3n) maxima, [3n..4n) population variances
 * about the FP64 mean, [4n] maximum row sum -- the reductions of examples/simulate.py:299-346
 * (net, mass, comp_min/max, comp_var, comp_fe_min).  Blocks until out is on the host. */
int pf5d_field_stats(int prec, const void* A, long ncell, long astride, int ncomp, double* out, void* stream);

/* Per-component FP64 dot products out[k] = sum_c A[c,k] B[c,k] of two contiguous (ncell, ncomp <= 8)
 * arrays: the sum of X . kappa lap X in the save-step gradient energy (examples/simulate.py:336-345).
 * Blocks until out is on the host. */
int pf5d_field_dot(int prec, const void* A, const void* B, long ncell, int ncomp, double* out, void* stream);

/* abs_max -- solver/operators.py:161-175.  Blocks until the value is on the host. */
int pf5d_abs_max(int prec, const void* a, long n, double* host_result, void* stream);

/* composition_change *= dt; current_composition += composition_change -- simulate.py:316-317 */
int pf5d_update(int prec, void* X, void* dX, long nelem, double dt, void* stream);

/* clamp_simplex -- solver/operators.py:262-317 */
int pf5d_clamp_simplex(int prec, void* X, long ncell, void* stream);

/* out_of_bounds statistics -- solver/system.py:12-53.  host_out5 =
 * [nonfinite count, max sum, min sum, min component, max component]; blocks. */
int pf5d_bounds_stats(int prec, const void* X, long ncell, double host_out5[5], void* stream);

/* pf5d_bounds_stats without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same kernel, same result bits.
 * The per-block partials live in a fixed 40 KiB device array (1024 blocks of five
 * doubles; the launch uses at most two blocks per CU) that each GPU holds until
 * the process ends; a device needing more blocks is rejected with PF5D_EINVAL.
 * Every call on a GPU shares that GPU's array, so calls that overlap on one GPU
 * (other host threads, other streams) race; serialise them.  Blocks. */
int pf5d_bounds_stats_noalloc(int prec, const void* X, long ncell, double host_out5[5], void* stream);

/* make_elastic_updater rank-1 path -- solver/elastic.py:147-263.  bhat_host:
 * kernel on the rfft half grid (k_shape) in J/m^3; l, X0: rank-1 vector and
 * reference composition.  apply(): mu[c*mstride + s] += psi[c] * l[s] with
 * psi = irfftn(bhat * rfftn(sum_s (X_s - X0_s) l_s)). */
int pf5d_elastic_create(int prec, int dim, const long* shape, const double* bhat_host, const double l[4],
                        const double X0[4], pf5d_elastic** out);
int pf5d_elastic_apply(pf5d_elastic* e, void* mu, long mstride, const void* X, void* stream);
/* psi of the last apply() (device pointer, ncell elements of prec) */
const void* pf5d_elastic_psi(const pf5d_elastic* e);
int pf5d_elastic_destroy(pf5d_elastic* e);

/* ---- device-resident fused step (examples/simulate.py:230-320 loop body) ---- */
typedef struct pf5d_sim_desc {
  int dim;            /* 1, 2, 3 */
  long shape[3];      /* first dim entries used */
  double cell_size;   /* h [m] */
  double kappa[4];    /* kappa_i * kappa_value */
  double X0[4];       /* reference composition (misfit origin) */
  double l[4];        /* rank-1 elastic vector; ignored if elastic == 0 */
  int elastic;        /* 1: bhat given */
  double ctld, dt_cfl, dt0;
  double lower_limit, upper_limit;
  double t_end;       /* stop when t >= t_end (inf: never) */
  int oob_interval;   /* bounds check period in steps (reference: 100; 0 = off) */
  double tol_soft, tol_hard;
} pf5d_sim_desc;

/* stream: NULL -> the engine creates its own non-blocking stream */
int pf5d_engine_create(int prec, void* stream, pf5d_engine** out);
int pf5d_engine_destroy(pf5d_engine* e);
/* X_host: (ncell,4) float64; bhat_host as for pf5d_elastic_create; returns slot.
 * A simulation holds at most 2^28 cells in FP32 and 2^27 in FP64 (PF5D_EINVAL
 * beyond: the step kernels use 32-bit offsets into its fields). */
int pf5d_engine_add(pf5d_engine* e, const pf5d_sim_desc* d, const double* X_host, const double* bhat_host,
                    const pf5d_table* tab, int* slot);
/* pf5d_engine_add with the initial field drawn on the device instead of by
 * apply_fluctuations on the host (solver/system.py:56-61, same distribution):
 * X = X0 + fluctuation * (noise - mean over the five components)[1:], noise
 * uniform in [-1, 1) from a counter-based generator (SplitMix64 of seed and the
 * value's index), so large fields need no host draw and no upload.  The same
 * seed gives the same FP64 field in either precision. */
int pf5d_engine_add_generated(pf5d_engine* e, const pf5d_sim_desc* d, unsigned long long seed, double fluctuation,
                              const double* bhat_host, const pf5d_table* tab, int* slot);
int pf5d_engine_remove(pf5d_engine* e, int slot);
/* K1 tile: variant index (pf5d_engine_num_variants), x chunk (3D), per dimension.
 * info (5*n ints, may be NULL), per variant: by, bz, nt of the potentials pass's
 * tiles (by == 0: two-pass form; by == -1: fused marching form, whose potentials
 * pass covers only the marching tiles' boundary rows and planes), then the flux pass:
 * cells per lane (0: per-plane tiles, > 0: marching along axis 0 in x chunks)
 * and its block size.  A variant that does not fit the simulations held at the
 * next step (the fused form: 3D, rows of 128 cells, axis 1 a multiple of 8,
 * axis 0 a multiple of the x chunk >= 3) runs the default tiling instead. */
int pf5d_engine_num_variants(pf5d_engine* e, int* n, int* info);
int pf5d_engine_set_tiling(pf5d_engine* e, int variant, int cx);
/* FFT passes: LDS bytes per work item in KiB (sets how many lines one block
 * transforms) and threads per block (128 or 256) */
int pf5d_engine_set_fft(pf5d_engine* e, int budget_kb, int nt);
/* FFT plane passes for 3D fields whose y and z lengths are a plane length (128, and 64 and 96 with 512 or
 * 1024 threads and 2 or 4 parts):
 * the z and y transforms of an x-plane in one work item, the plane split into
 * parts (2, 4 or 8) work items of nt threads (256, 512 or 1024); parts = 0: line
 * passes only (the default).  Where the plane passes need more LDS than a block
 * has, the line passes are used. */
int pf5d_engine_set_fft_planes(pf5d_engine* e, int parts, int nt);
/* Wave-resident FFT passes for 3D fields whose plane (y and z lengths) or x length
 * is 128: each 128-point line transformed in the registers of 8 threads of one
 * wavefront.  mode is a mask of the passes that use them: 1 the forward plane pass,
 * 2 the x pass, 4 the inverse plane pass (the others run the Stockham plane or line
 * passes above).  Defaults: 3 in FP32, 7 in FP64.  Every mode gives the same bits. */
int pf5d_engine_set_fft_wave(pf5d_engine* e, int mode);
/* enqueue nsteps (graph replay); does not block */
int pf5d_engine_step(pf5d_engine* e, int nsteps);
/* capture (without running them) the step graphs pf5d_engine_step(e, nsteps) replays, so
 * that the first such call does not pay for graph capture and instantiation (a timed
 * region, a latency-bound loop).  No counterpart in the reference (its loop launches
 * operators one by one). */
int pf5d_engine_prepare(pf5d_engine* e, int nsteps);
int pf5d_engine_sync(pf5d_engine* e);
/* state[16] = dt, t, last_xmax, mass[4], steps, status, smax, smin, cmin, cmax, nonfinite, mean0, mean1 */
int pf5d_engine_get_state(pf5d_engine* e, int slot, double* state16);
/* the states of slots 0..min(maxn, slot count)-1 in one device->host copy (16 doubles
 * each, layout as above; status 0 for free slots); *n = slots written */
int pf5d_engine_get_all_states(pf5d_engine* e, int maxn, double* out, int* n);
int pf5d_engine_set_time(pf5d_engine* e, int slot, double dt, double t);
/* copy fields to host as float64: which = 0 X, 1 dX/dt (of the last step), 2 psi */
int pf5d_engine_get_field(pf5d_engine* e, int slot, int which, double* host_out);
int pf5d_engine_set_field(pf5d_engine* e, int slot, const double* X_host);
/* device pointer of a field (which as above) for zero-copy consumers */
void* pf5d_engine_field_ptr(pf5d_engine* e, int slot, int which);
void* pf5d_engine_stream(pf5d_engine* e);
/* number of active slots and bytes of device memory held */
int pf5d_engine_stats(pf5d_engine* e, long* active, long* bytes);
/* launch configuration of the current work lists, for profiling reports: [K1 items, K2 items,
 * FFT items, K1 grid, 0 (dynamic K1 LDS: none), stencil variant, x chunk, FFT LDS budget bytes, FFT block
 * size, flux-pass grid, FFT plane parts in use (0: line passes only), FFT plane block size,
 * mask of the FFT passes running wave-resident (pf5d_engine_set_fft_wave)] */
int pf5d_engine_launch_info(pf5d_engine* e, long out[13]);
/* number of simulation groups the engine steps one after another (each group's per-step
 * working set fits the GPU's last-level cache; 1: all simulations step together) */
int pf5d_engine_num_groups(pf5d_engine* e, int* ngroups);

/* average device time of one launch of a stage (HIP events on the engine stream):
 * which = 0 stencil stage K1 (both of its passes), 1 the FFT passes, 2 dt controller, 3
 * update kernel in its statistics-only mode, 4 / 5 the potentials / flux pass of K1, 6
 * the FFT passes, K1 and the statistics-only update in sequence (K1 amid a step's traffic).  The timed launches rewrite the dX/dt and psi fields
 * from the current state; the call then restores what the next step reads (the
 * max|dX/dt| accumulator, the FP64 statistics), so the trajectory is unchanged. */
int pf5d_engine_time_kernel(pf5d_engine* e, int which, int reps, double* avg_ms);

/* per-stage device time of whole steps of a one-group engine with the marching two-pass
 * stencil: reps steps launched one by one on the engine stream with HIP events between
 * the stages.  out[6] = average ms of the FFT passes, the potentials pass, the flux pass
 * (its last block running the dt controller unless separate_dt), the separate dt launch
 * (0 when fused), the update pass and the whole step.  The steps advance the simulations
 * as pf5d_engine_step(e, reps) does.  A measurement helper with no counterpart in the
 * reference (solver/system.py steps without per-stage timing). */
int pf5d_engine_time_step_stages(pf5d_engine* e, int reps, int separate_dt, double* out);

#ifdef __cplusplus
}
#endif
#endif /* PF5D_H */
