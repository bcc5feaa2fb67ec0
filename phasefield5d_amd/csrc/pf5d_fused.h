// Device-resident fused explicit step: the types the step kernels share, and
// the stages after the stencil (the stencil stage K1 -- interpolation,
// kappa*lap(X) and elastic psi -> mu, Mob -> face fluxes -> dX/dt and the block
// max |dX/dt| -- is in pf5d_split.h).
//
//   DT  k_dtctl   : per simulation: max |dX/dt| -> adaptive dt (system.py:90-112),
//                   time, step count, FP64 mass / bounds statistics, status
//   K2  k_update  : X += dt dX/dt, simplex clamp, misfit s for the next FFT,
//                   FP64 partial sums (mass, bounds) per work item
//
// Every launch walks a work list that may span many simulations of different
// dimension and size (batched and serving modes use the same kernels).
//
// Canonical tile geometry: a simulation is viewed as (n0, n1, n2) with n2
// contiguous.  3D (NX,NY,NZ) marches along x (axis 0) with a (y,z) tile;
// 2D (NX,NY) is (1, NX, NY) with a single plane; 1D (N) is (1, 1, N).
// Axes present in the physical problem are summed in the reference order
// (slowest first), so the floating-point order of the Laplacian and of the
// divergence equals the reference's (operators.py:33-81, fluxes.py:296-395).
#pragma once
#include "pf5d_common.h"
#include "pf5d_flux.h"
#include "pf5d_interp.h"

namespace pf5d {

constexpr int NSTAT = 9;  // sum X (4), nonfinite, max sum, min sum, min comp, max comp

template <typename T>
struct SimDev {
  int dim;            // physical dimension 1..3
  int n0, n1, n2;     // canonical shape
  long ncell;
  T inv_dx, inv_dx2;  // 1/h, 1/h^2
  T kappa[4];         // kappa_i * kappa_value
  T l[4];             // rank-1 elastic vector (zeros when elasticity is off)
  T X0[4];            // misfit origin: X0 (FP64 mode), FP64 mean of the initial field (FP32 mode)
  int elastic;
  int use_mean;       // 1 when X0 holds the initial mean (the k=0 mode is removed before the FP32 FFT)
  int fresh;          // the field was just set: a prime launch writes its misfit s
  TableView<T> tab;
  T* X;
  Rem4* Xlo;          // FP32 mode: 24-bit rounding remainders of a cell's values (rem_decode; null in FP64 mode)
  T* dX;
  const T* psi;
  T* s;
  T* mu;              // two-pass stencil: mu (ncell x 4), Mob (ncell x 4 + ncell)
  T* mb;
  T* mb1;
  double ctld, dt_cfl, lower, upper, t_end;
  int oob_interval;   // bounds check every N steps (reference: 100)
  double tol_soft, tol_hard;
};

struct K1Item {
  int sim;
  int xa, xb;   // marching range along axis 0 (3D); 0..1 otherwise
  int y0, z0;   // tile origin on axes 1, 2
  int ty, tz;   // tile extent (<= BY, BZ)
};

struct K2Item {
  int sim;
  long c0, c1;  // cell range
};

// ---------------------------------------------------------------------------
// DT: one thread block per simulation
//   * reduce the previous K2 partials -> mass, bounds statistics (FP64, Neumaier)
//   * bounds check every oob_interval steps (system.py:12-53; simulate.py:231)
//   * adaptive dt (system.py:90-112) from max |dX/dt| of this step
// ---------------------------------------------------------------------------
PF5D_DEV double adaptive_dt(double dt, double x, double ctld, double t_cfl, double lo_lim, double hi_lim) {
  double raw;
  if (!(x > 0.0)) {
    raw = dt;
  } else {
    double lo = lo_lim / x, hi = hi_lim / x;
    if (lo > hi) {
      const double tmp = lo;
      lo = hi;
      hi = tmp;
    }
    raw = dt < lo ? lo : dt;   // numpy clip: min(max(dt, lo), hi)
    raw = raw > hi ? hi : raw;
  }
  double out = dt * 2.0;
  out = raw < out ? raw : out;
  const double cap = 0.005 * ctld;
  out = cap < out ? cap : out;
  out = t_cfl < out ? t_cfl : out;
  return out;
}

// The controller for one simulation, run by one block of NTB threads (red: an
// LDS array of NTB/64 rows): reduce the FP64 partials of the last update (mass,
// bounds), take max |dX/dt| (and reset it for the next step), then the bounds
// check, the end time and the adaptive dt (system.py:90-112).
template <typename T, int NTB>
PF5D_DEV void dt_control(int sim, const SimDev<T>* __restrict__ sims, SimState* __restrict__ st,
                         T* __restrict__ xmax_acc, const double* __restrict__ part,
                         const int* __restrict__ part_range, int advance, double (*red)[NSTAT], int tid) {
  const int p0 = part_range[2 * sim], p1 = part_range[2 * sim + 1];
  // The update pass writes the partials only in a prime launch and ahead of a
  // bounds check (when its step count is a multiple of the check interval); a
  // stepping controller sees that same count, so on other steps the partials
  // are the ones already reduced into the state, and are not read again
  // (uniform over the block: every thread reads the same state)
  const int iv = sims[sim].oob_interval;
  const bool reduce = !advance || (iv > 0 && (st[sim].steps % iv) == 0);
  double v[NSTAT];
  if (reduce) {  // (uniform over the block: every thread read the same state)
    // Neumaier sums of the mass partials, min/max of the rest
    double sum[4] = {0, 0, 0, 0}, cmp[4] = {0, 0, 0, 0};
    double nf = 0, smax = -1e300, smin = 1e300, cmin = 1e300, cmax = -1e300;
    for (int i = p0 + tid; i < p1; i += NTB) {
      const double* q = part + (long)i * NSTAT;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const double u = q[a], t = sum[a] + u;
        cmp[a] += (fabs(sum[a]) >= fabs(u)) ? ((sum[a] - t) + u) : ((u - t) + sum[a]);
        sum[a] = t;
      }
      nf += q[4];
      smax = fmax(smax, q[5]);
      smin = fmin(smin, q[6]);
      cmin = fmin(cmin, q[7]);
      cmax = fmax(cmax, q[8]);
    }
    const double v0[NSTAT] = {sum[0] + cmp[0], sum[1] + cmp[1], sum[2] + cmp[2], sum[3] + cmp[3], nf, smax, smin,
                              cmin, cmax};
#pragma unroll
    for (int k = 0; k < NSTAT; ++k) v[k] = v0[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
      for (int a = 0; a < 5; ++a) v[a] += __shfl_xor(v[a], off, 64);
      v[5] = fmax(v[5], __shfl_xor(v[5], off, 64));
      v[6] = fmin(v[6], __shfl_xor(v[6], off, 64));
      v[7] = fmin(v[7], __shfl_xor(v[7], off, 64));
      v[8] = fmax(v[8], __shfl_xor(v[8], off, 64));
    }
    if ((tid & 63) == 0)
      for (int k = 0; k < NSTAT; ++k) red[tid >> 6][k] = v[k];
  }
  // (and every thread has read the state before thread 0 changes it below)
  __syncthreads();
  if (tid == 0) {
    for (int w = 1; w < NTB / 64 && reduce; ++w) {
      for (int a = 0; a < 5; ++a) v[a] += red[w][a];
      v[5] = fmax(v[5], red[w][5]);
      v[6] = fmin(v[6], red[w][6]);
      v[7] = fmin(v[7], red[w][7]);
      v[8] = fmax(v[8], red[w][8]);
    }
    SimState& S = st[sim];
    const SimDev<T>& sd = sims[sim];
    // atomic exchange: the maxima were folded in with device-scope atomics
    const double x = (double)__hip_atomic_exchange(&xmax_acc[sim], T(0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    [&] {
      // statistics of the fields as the last update (or a prime launch) left
      // them: kept for inactive simulations only from a prime launch (advance
      // 0), which computes them from the current fields
      if (reduce && p1 > p0 && (!advance || (S.status & ST_ACTIVE))) {
        for (int a = 0; a < 4; ++a) {
          S.mass[a] = v[a];
          S.mean[a] = v[a] / (double)sd.ncell;
        }
        S.nonfinite = v[4];
        S.smax = v[5];
        S.smin = v[6];
        S.cmin = v[7];
        S.cmax = v[8];
      }
      if (!(S.status & ST_ACTIVE) || !advance) return;
      // bounds check of X_n before step n (the reference checks every 100 steps)
      if (sd.oob_interval > 0 && (S.steps % sd.oob_interval) == 0) {
        const double th = sd.tol_hard, ts = sd.tol_soft;
        const bool hard = S.nonfinite > 0 || S.smax > 1.0 + th || S.smin < -th || S.cmin < -th || S.cmax > 1.0 + th;
        const bool soft = S.smax > 1.0 + ts || S.smin < -ts || S.cmin < -ts || S.cmax > 1.0 + ts;
        if (soft) S.status |= ST_SOFT_OOB;
        if (hard) {
          S.status = (S.status & ~ST_ACTIVE) | ST_HARD_OOB;
          return;
        }
      }
      if (S.t >= sd.t_end) {
        S.status = (S.status & ~ST_ACTIVE) | ST_DONE;
        return;
      }
      const double dt = adaptive_dt(S.dt, x, sd.ctld, sd.dt_cfl, sd.lower, sd.upper);
      S.dt = dt;
      S.t = S.t + dt;
      S.last_xmax = x;
      S.steps += 1;
    }();
  }
  __syncthreads();  // red is reused by the next call
}

// one block per simulation: simulation list[blockIdx.x] (blockIdx.x without a list)
template <typename T, int NTB>
__global__ void __launch_bounds__(NTB) k_dtctl(const SimDev<T>* __restrict__ sims, SimState* __restrict__ st,
                                               T* __restrict__ xmax_acc, const double* __restrict__ part,
                                               const int* __restrict__ part_range, int nsims, int advance,
                                               const int* __restrict__ list) {
  if ((int)blockIdx.x >= nsims) return;
  const int sim = list ? list[blockIdx.x] : (int)blockIdx.x;
  __shared__ double red[NTB / 64][NSTAT];
  dt_control<T, NTB>(sim, sims, st, xmax_acc, part, part_range, advance, red, threadIdx.x);
}

// ---------------------------------------------------------------------------
// K2
// ---------------------------------------------------------------------------
template <typename T>
PF5D_DEV void neumaier_add(double& s, double& c, double v) {
  const double t = s + v;
  c += (fabs(s) >= fabs(v)) ? ((s - t) + v) : ((v - t) + s);
  s = t;
}

// mode 0: update (X += dt dX, clamp, misfit, partials); mode 1: prime (misfit and
// partials of the current X only)
template <typename T, int NT>
__global__ void __launch_bounds__(NT) k_update(const SimDev<T>* __restrict__ sims, const SimState* __restrict__ st,
                                               const K2Item* __restrict__ items, int nitems, double* __restrict__ part,
                                               int prime) {
  __shared__ double red[NT / 64][NSTAT];
  int ib, ie;
  block_range(nitems, ib, ie);
  for (int it = ib; it < ie; ++it) {
    const K2Item item = items[it];
    const SimState& S = st[item.sim];
    const bool run = prime || (S.status & ST_ACTIVE);
    const SimDev<T>& sd = sims[item.sim];
    const double dt = S.dt;
    // the FP64 statistics (mass, bounds) are read on the device only by the
    // bounds check of the next step, every oob_interval steps (S.steps already
    // counts this step); the host refreshes them with a prime launch before it
    // reads a state, so other steps skip them
    const bool stats = prime || (sd.oob_interval > 0 && (S.steps % sd.oob_interval) == 0);
    double sum[4] = {0, 0, 0, 0}, cmp[4] = {0, 0, 0, 0};
    double nf = 0, smax = -1e300, smin = 1e300, cmin = 1e300, cmax = -1e300;
    T ref[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) ref[a] = sd.X0[a];
#pragma unroll 2
    for (long c = item.c0 + threadIdx.x; c < item.c1; c += NT) {
      V4<T> x = ld4g<T>(sd.X + 4 * c);
      double xd[4];  // the state in FP64 (FP32 mode: X with its remainder)
      if (sd.Xlo) {
        int lo[4];
        rem_unpack(ld_rem(sd.Xlo + c), lo);
#pragma unroll
        for (int a = 0; a < 4; ++a) xd[a] = rem_decode((float)x.v[a], lo[a]);
      } else {
#pragma unroll
        for (int a = 0; a < 4; ++a) xd[a] = (double)x.v[a];
      }
      if (run && !prime) {
        const V4<T> d = ld4g<T>(sd.dX + 4 * c);
        if (!sd.Xlo) {
          // FP64: composition_change *= dt; X += composition_change; clamp (simulate.py:316-320)
#pragma unroll
          for (int a = 0; a < 4; ++a) x.v[a] = x.v[a] + (T)((double)d.v[a] * dt);
          const T tot = ((x.v[0] + x.v[1]) + x.v[2]) + x.v[3];
          if (tot > T(1)) {
            const T inv = T(1) / tot;
#pragma unroll
            for (int a = 0; a < 4; ++a) x.v[a] = x.v[a] * inv;
          }
#pragma unroll
          for (int a = 0; a < 4; ++a) xd[a] = (double)x.v[a];
          st4g<T>(sd.X + 4 * c, x);
        } else {
          // FP32 state with a 24-bit remainder: update and clamp in FP64, then split
          // into the nearest FP32 value and what it leaves out, so increments below
          // half an FP32 ulp accumulate instead of being rounded away
#pragma unroll
          for (int a = 0; a < 4; ++a) xd[a] = xd[a] + (double)d.v[a] * dt;
          const double tot = ((xd[0] + xd[1]) + xd[2]) + xd[3];
          if (tot > 1.0) {
            const double inv = 1.0 / tot;
#pragma unroll
            for (int a = 0; a < 4; ++a) xd[a] = xd[a] * inv;
          }
          int lo[4];
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            x.v[a] = (T)xd[a];
            lo[a] = rem_encode(xd[a], (float)x.v[a]);
          }
          st4g<T>(sd.X + 4 * c, x);
          st_rem(sd.Xlo + c, rem_pack(lo));
        }
      }
      // the misfit s: from the updated state in a step; a prime launch writes it
      // only for a field just set, and otherwise keeps the one the last step
      // wrote (X + X_lo recombined can differ from the updated FP64 value
      // below FP32 resolution, so recomputing it could round differently)
      if (sd.elastic && run && (!prime || sd.fresh)) {
        T acc = (T)(xd[0] - (double)ref[0]) * sd.l[0];
#pragma unroll
        for (int a = 1; a < 4; ++a) acc = mad<T>((T)(xd[a] - (double)ref[a]), sd.l[a], acc);
        gst<T>(sd.s + c, acc);
      }
      if (stats) {
        double csum = 0;
        bool fin = true;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const double v = xd[a];
          neumaier_add<T>(sum[a], cmp[a], v);
          fin = fin && isfinite(v);
          csum += v;
          cmin = fmin(cmin, v);
          cmax = fmax(cmax, v);
        }
        if (!fin) nf += 1;
        smax = fmax(smax, csum);
        smin = fmin(smin, csum);
      }
    }
    if (!stats) continue;  // uniform over the block: the item's simulation
    double v[NSTAT] = {sum[0] + cmp[0], sum[1] + cmp[1], sum[2] + cmp[2], sum[3] + cmp[3], nf, smax, smin, cmin, cmax};
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
      for (int a = 0; a < 5; ++a) v[a] += __shfl_xor(v[a], off, 64);
      v[5] = fmax(v[5], __shfl_xor(v[5], off, 64));
      v[6] = fmin(v[6], __shfl_xor(v[6], off, 64));
      v[7] = fmin(v[7], __shfl_xor(v[7], off, 64));
      v[8] = fmax(v[8], __shfl_xor(v[8], off, 64));
    }
    if ((threadIdx.x & 63) == 0)
      for (int k = 0; k < NSTAT; ++k) red[threadIdx.x >> 6][k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int w = 1; w < NT / 64; ++w) {
        for (int a = 0; a < 5; ++a) v[a] += red[w][a];
        v[5] = fmax(v[5], red[w][5]);
        v[6] = fmin(v[6], red[w][6]);
        v[7] = fmin(v[7], red[w][7]);
        v[8] = fmax(v[8], red[w][8]);
      }
      for (int k = 0; k < NSTAT; ++k) part[(long)it * NSTAT + k] = v[k];
    }
    __syncthreads();
  }
}

}  // namespace pf5d
