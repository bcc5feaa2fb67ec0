// Host side of the phasefield5d MI355X step library: C ABI, FFT plans, the
// device-resident engine (work lists over many simulations, graph replay).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pf5d.h"
#include "pf5d_common.h"
#include "pf5d_fft.h"
#include "pf5d_flux.h"
#include "pf5d_fused.h"
#include "pf5d_interp.h"
#include "pf5d_split.h"
#include "pf5d_instances.h"

namespace pf5d {
PF5D_FFT32_INSTANCES
PF5D_FFT64_INSTANCES
PF5D_STENCIL_INSTANCES
}  // namespace pf5d

using namespace pf5d;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

struct Pf5dError : std::runtime_error {
  int code;
  Pf5dError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                   \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess)                                                                           \
      throw Pf5dError(PF5D_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_) + " at line " + \
                                        std::to_string(__LINE__));                                  \
  } while (0)

static void require(bool ok, const std::string& msg, int code = PF5D_EINVAL) {
  if (!ok) throw Pf5dError(code, msg);
}

template <typename F>
static int guarded(F&& f) {
  try {
    f();
    return PF5D_OK;
  } catch (const Pf5dError& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return PF5D_ENOMEM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return PF5D_EDEVICE;
  }
}

static void* dmalloc(size_t bytes) {
  void* p = nullptr;
  if (bytes == 0) bytes = 16;
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess) throw Pf5dError(PF5D_ENOMEM, "hipMalloc(" + std::to_string(bytes) + "): " + hipGetErrorString(e));
  return p;
}
static void dfree(void* p) {
  if (p) (void)hipFree(p);
}

struct DevInfo {
  int cus = 256, lds = 65536, warp = 64;
};
static DevInfo& devinfo() {
  static DevInfo d;
  static bool init = false;
  if (!init) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIPCHK(hipGetDeviceProperties(&p, dev));
    d.cus = p.multiProcessorCount;
    d.lds = (int)p.sharedMemPerBlock;
    d.warp = p.warpSize;
    init = true;
  }
  return d;
}

static inline unsigned nblocks(long n, int bs) { return (unsigned)std::max<long>(1, (n + bs - 1) / bs); }

// grid size rounded up to a multiple of the XCD count, so that block_range
// gives each XCD one contiguous share of a work list (extra blocks get none)
static unsigned xcd_grid(long g) {
  g = std::max<long>(1, g);
  return (unsigned)((g + NXCD - 1) / NXCD * NXCD);
}

// library-internal grow-only device scratch for reductions
struct Scratch {
  void* p = nullptr;
  size_t bytes = 0;
  void* get(size_t b) {
    if (b > bytes) {
      dfree(p);
      p = dmalloc(b);
      bytes = b;
    }
    return p;
  }
};
static Scratch& grow_scratch() {
  static thread_local Scratch s;
  return s;
}

// ---------------------------------------------------------------------------
// tables
// ---------------------------------------------------------------------------
struct pf5d_table {
  int prec, num;
  double resolution;
  void* data;
  void* cubes;  // cube table of the linear region (null when num < 5)
  int cube_s;
  bool complete;  // every node of the simplex holds a row (no NaN from a missing one)
};

template <typename T>
static TableView<T> view(const pf5d_table* t) {
  TableView<T> v;
  v.data = reinterpret_cast<const T*>(t->data);
  v.num = t->num;
  v.inv_step = (T)(1.0 / t->resolution);
  v.thresh = (T)(1.0 - 4.0 * t->resolution - 1e-9);
  v.step = t->resolution;
  v.cubes = reinterpret_cast<const T*>(t->cubes);
  v.cube_s = t->cube_s;
  return v;
}

// one thread per dense cube position; cubes of the linear region copy their 16
// corners into the cube table's order
template <typename T>
__global__ void k_build_cubes(const T* __restrict__ dense, int num, int S, T* __restrict__ cubes) {
  constexpr int NP = TablePad<T>::value, VN = Vec16<T>::n;
  const long id = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long m = num - 1;
  if (id >= m * m * m * m) return;
  const int i3 = (int)(id % m), i2 = (int)(id / m % m), i1 = (int)(id / (m * m) % m), i0 = (int)(id / (m * m * m));
  if (i0 + i1 + i2 + i3 > S) return;
  T* out = cubes + simplex_rank4(i0, i1, i2, i3, S) * CUBE_VALUES;
  for (int b = 0; b < 16; ++b) {
    const long row = (((long)(i0 + (b >> 3)) * num + i1 + ((b >> 2) & 1)) * num + i2 + ((b >> 1) & 1)) * num +
                     i3 + (b & 1);
    const T* node = dense + row * NP;
    for (int g = 0; g < 8 / VN; ++g)
      for (int e = 0; e < VN; ++e) out[(g * 16 + b) * VN + e] = node[g * VN + e];
    out[128 + b] = node[8];
  }
}

template <typename T>
static void upload_table(pf5d_table* t, const double* grid, int nprops) {
  constexpr int NP = TablePad<T>::value;
  const long nn = (long)t->num * t->num * t->num * t->num;
  std::vector<T> h((size_t)nn * NP, T(0));
  for (long i = 0; i < nn; ++i)
    for (int p = 0; p < nprops; ++p) h[i * NP + p] = (T)grid[i * nprops + p];
  t->complete = true;
  const int n = t->num;
  for (int a = 0; a < n && t->complete; ++a)
    for (int b = 0; a + b < n && t->complete; ++b)
      for (int c = 0; a + b + c < n && t->complete; ++c)
        for (int d = 0; a + b + c + d < n; ++d)
          if (std::isnan(grid[((((long)a * n + b) * n + c) * n + d) * nprops])) {
            t->complete = false;
            break;
          }
  t->data = dmalloc(h.size() * sizeof(T));
  HIPCHK(hipMemcpy(t->data, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  // the cube table: lower corners with index sum <= num - 5 hold every point
  // of the linear region (sum X <= 1 - 4h - 1e-9)
  t->cube_s = t->num - 5;
  t->cubes = nullptr;
  // (cube_rank's 32-bit arithmetic holds up to CUBE_MAX_S; a finer table, whose
  // cube table would pass half a terabyte, interpolates from the dense rows)
  const long ncubes = t->cube_s <= CUBE_MAX_S ? simplex_count4(t->cube_s) : 0;
  if (ncubes > 0) {
    t->cubes = dmalloc((size_t)ncubes * CUBE_VALUES * sizeof(T));
    const long m = t->num - 1, npos = m * m * m * m;
    hipLaunchKernelGGL(k_build_cubes<T>, dim3((unsigned)((npos + 255) / 256)), dim3(256), 0, nullptr,
                       reinterpret_cast<const T*>(t->data), t->num, t->cube_s, reinterpret_cast<T*>(t->cubes));
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(nullptr));
  }
}

// ---------------------------------------------------------------------------
// FFT plans and batched pencil passes
// ---------------------------------------------------------------------------
// the FFT plan of a line length (the radix order the kernels assume)
static std::vector<int> factorize(int n) {
  std::vector<int> r;
  for (int p = 0; fft_radix(n, p) != 0; ++p) r.push_back(fft_radix(n, p));
  return r;
}

// exp(-2 pi i m / N) with integer reduction and long-double evaluation
static void root_of_unity(long m, long N, double& re, double& im) {
  m %= N;
  const long double a = -2.0L * 3.141592653589793238462643383279502884L * (long double)m / (long double)N;
  re = (double)cosl(a);
  im = (double)sinl(a);
}

template <typename T>
struct FftPlans {
  std::map<int, int> index;  // n -> plan id
  std::vector<LinePlanDev> plans;
  std::vector<cx<T>> tw, roots;
  LinePlanDev* d_plans = nullptr;
  cx<T>* d_tw = nullptr;
  cx<T>* d_roots = nullptr;
  bool dirty = false;

  int get(int n) {
    auto it = index.find(n);
    if (it != index.end()) return it->second;
    std::vector<int> rad = factorize(n);
    require((int)rad.size() <= FFT_MAXPASS, "FFT length has too many factors: " + std::to_string(n));
    LinePlanDev pl{};
    pl.n = n;
    pl.npass = (int)rad.size();
    long Ns = 1;
    for (int p = 0; p < pl.npass; ++p) {
      const int R = rad[p];
      pl.radix[p] = R;
      pl.tw_off[p] = (int)tw.size();
      for (long k = 0; k < Ns; ++k)
        for (int r = 1; r < R; ++r) {
          double re, im;
          root_of_unity((long)r * k, Ns * R, re, im);
          tw.push_back({(T)re, (T)im});
        }
      pl.root_off[p] = (int)roots.size();
      if (fft_generic_radix(R))
        for (int m = 0; m < R; ++m) {
          double re, im;
          root_of_unity(m, R, re, im);
          roots.push_back({(T)re, (T)im});
        }
      Ns *= R;
    }
    pl.nroot = (int)roots.size() - pl.root_off[0];
    require(pl.n + pl.nroot == fft_head_slots(n), "FFT plan roots do not match the LDS layout");
    // a line whose twiddles, roots and two buffers exceed a block's LDS reads
    // twiddles and roots from global memory (1D lines of a few thousand points)
    pl.tw_global = ((size_t)fft_head_slots(n) + 2 * (size_t)line_stride(n)) * sizeof(cx<T>) > (size_t)devinfo().lds;
    require(2 * (size_t)line_stride(n) * sizeof(cx<T>) <= (size_t)devinfo().lds,
            "FFT line too long for LDS: " + std::to_string(n) + " points");
    int id = (int)plans.size();
    plans.push_back(pl);
    index[n] = id;
    dirty = true;
    return id;
  }
  void upload() {
    if (!dirty) return;
    dfree(d_plans);
    dfree(d_tw);
    dfree(d_roots);
    d_plans = (LinePlanDev*)dmalloc(plans.size() * sizeof(LinePlanDev));
    d_tw = (cx<T>*)dmalloc(std::max<size_t>(1, tw.size()) * sizeof(cx<T>));
    d_roots = (cx<T>*)dmalloc(std::max<size_t>(1, roots.size()) * sizeof(cx<T>));
    HIPCHK(hipMemcpy(d_plans, plans.data(), plans.size() * sizeof(LinePlanDev), hipMemcpyHostToDevice));
    if (!tw.empty()) HIPCHK(hipMemcpy(d_tw, tw.data(), tw.size() * sizeof(cx<T>), hipMemcpyHostToDevice));
    if (!roots.empty())
      HIPCHK(hipMemcpy(d_roots, roots.data(), roots.size() * sizeof(cx<T>), hipMemcpyHostToDevice));
    dirty = false;
  }
  ~FftPlans() {
    dfree(d_plans);
    dfree(d_tw);
    dfree(d_roots);
  }
};

// Rows of the half spectrum (2D and 3D) are padded to whole 128-byte lines, so
// that groups of lines along the strided axes start on cache-line boundaries
// and neighbouring work items never share a line (the padding is not read).
template <typename T>
static long spec_row(int dim, const long* shape) {
  const long nh = shape[dim - 1] / 2 + 1;
  if (dim == 1) return nh;
  const long a = 128 / (long)sizeof(cx<T>);
  return (nh + a - 1) / a * a;
}
template <typename T>
static long spectrum_size(int dim, const long* shape) {
  long rows = 1;
  for (int a = 0; a < dim - 1; ++a) rows *= shape[a];
  return rows * spec_row<T>(dim, shape);
}

// plane passes: kernel and LDS bytes for a plane length and a part count
template <typename T>
struct PlaneKernels {
  using Fn = void (*)(const FftPlaneItem*, int, const LinePlanDev*, const cx<T>*, const SimState*);
  Fn fwd = nullptr, inv = nullptr;
  size_t fwd_lds = 0, inv_lds = 0;
  int nt = 0;
};
template <typename T, int NT, int N, int P>
static PlaneKernels<T> plane_kernels_np() {
  return {&k_fftp_fwd<T, NT, N, P>, &k_fftp_inv<T, NT, N, P>, PlaneFwd<T, N, P>::SLOTS * sizeof(cx<T>),
          PlaneInv<T, N, P>::SLOTS * sizeof(cx<T>)};
}
template <typename T, int NT>
static PlaneKernels<T> plane_kernels_nt(int n, int parts) {
  // (128-point FP64 planes are not built for 1024 threads: held to 128 VGPRs,
  // those kernels spill)
  if constexpr (!(sizeof(T) == 8 && NT == 1024)) {
    if (n == 128 && parts == 2) return plane_kernels_np<T, NT, 128, 2>();
    if (n == 128 && parts == 4) return plane_kernels_np<T, NT, 128, 4>();
    if (n == 128 && parts == 8) return plane_kernels_np<T, NT, 128, 8>();
  }
  if constexpr (NT != 256) {
    if (n == 64 && parts == 2) return plane_kernels_np<T, NT, 64, 2>();
    if (n == 64 && parts == 4) return plane_kernels_np<T, NT, 64, 4>();
    if (n == 96 && parts == 2) return plane_kernels_np<T, NT, 96, 2>();
    if (n == 96 && parts == 4) return plane_kernels_np<T, NT, 96, 4>();
  }
  return {};
}
// the block sizes, plane lengths and part counts built; fwd == nullptr when there is none
template <typename T>
static PlaneKernels<T> plane_kernels(int n, int parts, int nt) {
  if (nt == 256) return plane_kernels_nt<T, 256>(n, parts);
  if (nt == 512) return plane_kernels_nt<T, 512>(n, parts);
  if (nt == 1024) return plane_kernels_nt<T, 1024>(n, parts);
  return {};
}

// wave-resident passes of 128-point planes and x lines (pf5d_fftreg.h), one
// set per precision; the engine's wave mode is a mask of the passes that use
// them: 1 the forward plane pass, 2 the x pass, 4 the inverse plane pass
constexpr int FFT_WAVE_ALL = 7;
template <typename T>
struct WaveKernels {
  using PFn = void (*)(const FftPlaneItem*, int, const LinePlanDev*, const cx<T>*, const SimState*);
  using XFn = void (*)(const FrXItem*, int, const LinePlanDev*, const cx<T>*, const SimState*);
  PFn fwd, inv;
  size_t fwd_lds, inv_lds;
  int fwd_nt, inv_nt, fwd_parts, inv_parts;
  XFn x;
  size_t x_lds;
  int x_nt;
};
template <typename T>
static WaveKernels<T> wave_kernels() {
  // FP32: forward planes in blocks of 8 waves, inverse planes of 9 (all 65
  // spectrum columns in one round), 2 parts each; FP64 (twice the LDS per line):
  // 4 waves, 2 parts.  x lines in blocks of one wave: at 128^3 the pass has about
  // one wave per SIMD, and one-wave blocks spread it over every CU (FP32 x pass
  // 14.5 us in four-wave blocks, 10.3 us in one-wave blocks; two-wave blocks 11.2)
  if constexpr (sizeof(T) == 4)
    return {&k_frp_fwd<T, 8, 2>, &k_frp_inv<T, 9, 2>, FrPlaneFwd<T, 8, 2>::BYTES, FrPlaneInv<T, 9, 2>::BYTES,
            8 * 64, 9 * 64, 2, 2, &k_frx<T, 1>, FrX<T, 1>::BYTES, 1 * 64};
  else
    return {&k_frp_fwd<T, 4, 2>, &k_frp_inv<T, 4, 2>, FrPlaneFwd<T, 4, 2>::BYTES, FrPlaneInv<T, 4, 2>::BYTES,
            4 * 64, 4 * 64, 2, 2, &k_frx<T, 1>, FrX<T, 1>::BYTES, 1 * 64};
}

// the five stages of psi = irfftn(Bhat rfftn(s)) and their work items
template <typename T>
struct FftBatch {
  // plans are owned per batch: their device arrays are baked into captured
  // graphs, so they may only be reallocated when the owner rebuilds its graphs
  FftPlans<T> plans;
  std::vector<FftItem> stage[5];  // 0: rows R2C / 1D line, 1: y fwd, 2: x conv, 3: y inv, 4: rows C2R
  size_t lds[5] = {0, 0, 0, 0, 0};
  FftItem* d_items = nullptr;
  int offs[6] = {0, 0, 0, 0, 0, 0};
  // tunables: LDS bytes per work item (sets lines per item) and threads per block
  size_t budget = 16 * 1024;
  int nt = 128;
  // plane passes (3D fields whose y and z lengths are a plane length): parts
  // per plane, 0 = line passes only
  int parts = 0, pnt = 256;  // parts per plane (0 = line passes only), threads per block
  // plane items per plane length (one launch per length and direction)
  struct PlaneGroup {
    // per direction: kernel, LDS bytes, threads per block
    typename WaveKernels<T>::PFn kf = nullptr, ki = nullptr;
    size_t lf = 0, li = 0;
    int nf = 0, ni = 0;
    std::vector<FftPlaneItem> fwd, inv;
    int off = 0;  // first item in d_pitems (forward items, then inverse items)
  };
  std::map<int, PlaneGroup> pgroups;
  FftPlaneItem* d_pitems = nullptr;
  bool has_planes() const { return !pgroups.empty(); }
  // wave-resident passes (mask, FFT_WAVE_ALL), where lengths are 128
  int wave = 0;
  int wave_used = 0;  // the passes that use them in the current items
  int parts_used = 0;  // parts per plane of the forward plane passes in use
  std::vector<FrXItem> xitems;
  FrXItem* d_xitems = nullptr;

  static bool paired(int kind) { return kind == FK_R2C || kind == FK_C2R; }
  // complex lines per item: two line buffers within the LDS budget (real-line
  // stages carry two lines per complex line)
  // LDS slots ahead of the line buffers of length n's items (none when the
  // plan keeps its twiddles in global memory)
  size_t head_slots(int n) const {
    const auto it = plans.index.find(n);
    return (it != plans.index.end() && plans.plans[it->second].tw_global) ? 0 : (size_t)fft_head_slots(n);
  }
  // LDS bytes per complex line of an item: two line buffers, and the Bhat
  // values CONV items of the compile-time lengths gather next to them
  static size_t line_bytes(int n, int kind) {
    return 2 * (size_t)line_stride(n) * sizeof(cx<T>) +
           ((fft_ct_length(n) && kind == FK_CONV) ? (size_t)n * sizeof(T) : 0);
  }
  int pick_g(int n, long nlines, int kind) const {
    const size_t head = head_slots(n) * sizeof(cx<T>);
    const size_t avail = budget > head ? budget - head : 0;
    long g = (long)(avail / line_bytes(n, kind));
    g = std::max<long>(1, std::min<long>(g, 128));
    if (paired(kind)) g *= 2;
    return (int)std::min<long>(g, nlines);
  }
  void push(int st, FftItem it, int n) {
    const size_t gl = paired(it.kind) ? (it.G + 1) / 2 : it.G;
    stage[st].push_back(it);
    lds[st] = std::max(lds[st], head_slots(n) * sizeof(cx<T>) + gl * line_bytes(n, it.kind));
  }
  // lines [0, nlines) with line l at base + l*ls; split into groups of G (a
  // multiple of align lines when G >= align)
  void add_lines(int st, FftItem proto, int n, long nlines, long line_unit_in, long line_unit_out, long line_unit_b,
                 int align = 1) {
    int G = pick_g(n, nlines, proto.kind);
    if (G >= align) G -= G % align;
    for (long l0 = 0; l0 < nlines; l0 += G) {
      FftItem it = proto;
      it.G = (int)std::min<long>(G, nlines - l0);
      it.in = (const char*)proto.in + l0 * line_unit_in;
      it.out = (char*)proto.out + l0 * line_unit_out;
      it.bhat = proto.bhat ? (const char*)proto.bhat + l0 * line_unit_b : nullptr;
      push(st, it, n);
    }
  }
  // s (real, ncell) -> psi (real) through spectrum C (complex, rows of nh) and bhat (real, C layout)
  void add_sim(int sim, int dim, const long* shape, const T* s, cx<T>* C, T* psi,
               const T* bhat) {
    const int nl = (int)shape[dim - 1];
    const long nh = nl / 2 + 1, nhp = spec_row<T>(dim, shape);
    const int al = 128 / (int)sizeof(cx<T>);  // lines per 128-byte segment along a strided axis
    const long rows = [&] {
      long r = 1;
      for (int a = 0; a < dim - 1; ++a) r *= shape[a];
      return r;
    }();
    FftItem b{};
    b.sim = sim;
    if (dim == 1) {
      b.kind = FK_LINE;
      b.plan = plans.get(nl);
      b.G = 1;
      b.in = s;
      b.out = psi;
      b.bhat = bhat;
      b.ps_in = 1;
      b.ls_in = nl;
      b.ps_out = 1;
      b.ls_out = nl;
      push(0, b, nl);
      return;
    }
    // rows: R2C
    b.kind = FK_R2C;
    b.plan = plans.get(nl);
    b.in = s;
    b.out = C;
    b.ps_in = 1;
    b.ls_in = nl;
    b.ps_out = 1;
    b.ls_out = nhp;
    size_t n0 = stage[0].size();
    add_lines(0, b, nl, rows, sizeof(T) * nl, sizeof(cx<T>) * nhp, 0);
    const size_t nrows_r2c = stage[0].size() - n0;
    // rows: C2R
    b.kind = FK_C2R;
    b.in = C;
    b.out = psi;
    b.ps_in = 1;
    b.ls_in = nhp;
    b.ps_out = 1;
    b.ls_out = nl;
    n0 = stage[4].size();
    add_lines(4, b, nl, rows, sizeof(cx<T>) * nhp, sizeof(T) * nl, 0);
    const size_t nrows_c2r = stage[4].size() - n0;
    if (dim == 2) {
      const int nx = (int)shape[0];
      b.kind = FK_CONV;
      b.plan = plans.get(nx);
      b.in = C;
      b.out = C;
      b.bhat = bhat;
      b.ps_in = b.ps_out = b.ps_b = nhp;
      b.ls_in = b.ls_out = b.ls_b = 1;
      add_lines(2, b, nx, nh, sizeof(cx<T>), sizeof(cx<T>), sizeof(T), al);
      return;
    }
    const int nx = (int)shape[0], ny = (int)shape[1];
    const size_t lmax = (size_t)devinfo().lds;
    const WaveKernels<T> wk = wave_kernels<T>();
    // per direction: the wave-resident plane pass, else the Stockham plane pass
    // (parts > 0), else the line passes (rows, then y lines)
    const bool wave_len = (ny == FR_N && nl == FR_N);
    PlaneKernels<T> ok{};
    if (ny == nl && parts > 0) ok = plane_kernels<T>(nl, parts, pnt);
    const bool wf = (wave & 1) && wave_len && wk.fwd_lds <= lmax;
    const bool wi = (wave & 4) && wave_len && wk.inv_lds <= lmax;
    const bool of = !wf && ok.fwd && ok.fwd_lds <= lmax && ok.inv_lds <= lmax;
    const bool oi = !wi && ok.inv && ok.fwd_lds <= lmax && ok.inv_lds <= lmax;
    if (wf || of || wi || oi) {
      PlaneGroup& pg = pgroups[nl];
      const int plan = plans.get(nl);
      if (wf || of) {
        // rows and y lines of each plane in the forward plane pass; its row stage above is dropped
        stage[0].resize(stage[0].size() - nrows_r2c);
        const auto kf = wf ? wk.fwd : ok.fwd;
        require(pg.fwd.empty() || pg.kf == kf, "FFT plane passes of one length must share their kernels");
        pg.kf = kf;
        pg.lf = wf ? wk.fwd_lds : ok.fwd_lds;
        pg.nf = wf ? wk.fwd_nt : pnt;
        const int pf = wf ? wk.fwd_parts : parts;
        for (int x = 0; x < nx; ++x)
          for (int p = 0; p < pf; ++p) pg.fwd.push_back({sim, plan, x, p, s, C, (int)nhp});
        parts_used = pf;
        if (wf) wave_used |= 1;
      }
      if (wi || oi) {
        stage[4].resize(stage[4].size() - nrows_c2r);
        const auto ki = wi ? wk.inv : ok.inv;
        require(pg.inv.empty() || pg.ki == ki, "FFT plane passes of one length must share their kernels");
        pg.ki = ki;
        pg.li = wi ? wk.inv_lds : ok.inv_lds;
        pg.ni = wi ? wk.inv_nt : pnt;
        const int pi = wi ? wk.inv_parts : parts;
        for (int x = 0; x < nx; ++x)
          for (int p = 0; p < pi; ++p) pg.inv.push_back({sim, plan, x, p, C, psi, (int)nhp});
        if (wi) wave_used |= 4;
      }
    }
    // y lines of the directions without a plane pass: for each x plane, lines kz in
    // [0, nh) with point stride nhp
    for (int x = 0; x < nx; ++x) {
      FftItem y{};
      y.sim = sim;
      y.plan = plans.get(ny);
      y.in = C + (long)x * ny * nhp;
      y.out = C + (long)x * ny * nhp;
      y.ps_in = y.ps_out = nhp;
      y.ls_in = y.ls_out = 1;
      if (!(wf || of)) {
        y.kind = FK_FWD;
        add_lines(1, y, ny, nh, sizeof(cx<T>), sizeof(cx<T>), 0, al);
      }
      if (!(wi || oi)) {
        y.kind = FK_INV;
        add_lines(3, y, ny, nh, sizeof(cx<T>), sizeof(cx<T>), 0, al);
      }
    }
    if ((wave & 2) && nx == FR_N && wk.x_lds <= lmax) {
      // x pass, wave-resident: items of (waves per block) x 8 lines (ky, kz), kz fastest
      const int plan = plans.get(nx);
      const int per = wk.x_nt / 8;
      const long nlines = (long)ny * nh;
      for (long l0 = 0; l0 < nlines; l0 += per)
        xitems.push_back({sim, plan, (int)l0, (int)std::min<long>(per, nlines - l0), ny, (int)nh, (int)nhp, C,
                          bhat});
      wave_used |= 2;
      return;
    }
    // x pass with Bhat: for each y, lines kz in [0, nh) with point stride ny*nhp
    for (int y = 0; y < ny; ++y) {
      FftItem c = b;
      c.kind = FK_CONV;
      c.plan = plans.get(nx);
      c.in = C + (long)y * nhp;
      c.out = C + (long)y * nhp;
      c.bhat = bhat + (long)y * nhp;
      c.ps_in = c.ps_out = c.ps_b = (long)ny * nhp;
      c.ls_in = c.ls_out = c.ls_b = 1;
      add_lines(2, c, nx, nh, sizeof(cx<T>), sizeof(cx<T>), sizeof(T), al);
    }
  }
  void clear() {
    for (auto& s : stage) s.clear();
    for (auto& l : lds) l = 0;
    pgroups.clear();
    xitems.clear();
    wave_used = 0;
    parts_used = 0;
  }
  void upload() {
    dfree(d_items);
    d_items = nullptr;
    dfree(d_pitems);
    d_pitems = nullptr;
    dfree(d_xitems);
    d_xitems = nullptr;
    if (!xitems.empty()) {
      d_xitems = (FrXItem*)dmalloc(xitems.size() * sizeof(FrXItem));
      HIPCHK(hipMemcpy(d_xitems, xitems.data(), xitems.size() * sizeof(FrXItem), hipMemcpyHostToDevice));
    }
    if (has_planes()) {
      std::vector<FftPlaneItem> pall;
      for (auto& kv : pgroups) {
        kv.second.off = (int)pall.size();
        pall.insert(pall.end(), kv.second.fwd.begin(), kv.second.fwd.end());
        pall.insert(pall.end(), kv.second.inv.begin(), kv.second.inv.end());
      }
      d_pitems = (FftPlaneItem*)dmalloc(pall.size() * sizeof(FftPlaneItem));
      HIPCHK(hipMemcpy(d_pitems, pall.data(), pall.size() * sizeof(FftPlaneItem), hipMemcpyHostToDevice));
    }
    std::vector<FftItem> all;
    for (int i = 0; i < 5; ++i) {
      offs[i] = (int)all.size();
      all.insert(all.end(), stage[i].begin(), stage[i].end());
    }
    offs[5] = (int)all.size();
    if (all.empty()) return;
    d_items = (FftItem*)dmalloc(all.size() * sizeof(FftItem));
    HIPCHK(hipMemcpy(d_items, all.data(), all.size() * sizeof(FftItem), hipMemcpyHostToDevice));
  }
  void run_planes(hipStream_t stream, const SimState* st, bool inverse) {
    const DevInfo& di = devinfo();
    for (auto& kv : pgroups) {
      const PlaneGroup& pg = kv.second;
      const int nf = (int)pg.fwd.size(), n = inverse ? (int)pg.inv.size() : nf;
      if (n == 0) continue;
      const unsigned grid = xcd_grid(std::min<long>(n, (long)di.cus * 8));
      hipLaunchKernelGGL(inverse ? pg.ki : pg.kf, dim3(grid), dim3(inverse ? pg.ni : pg.nf),
                         inverse ? pg.li : pg.lf, stream, d_pitems + pg.off + (inverse ? nf : 0), n, plans.d_plans,
                         plans.d_tw, st);
      HIPCHK(hipGetLastError());
    }
  }
  void run_x(hipStream_t stream, const SimState* st) {
    if (xitems.empty()) return;
    const WaveKernels<T> wk = wave_kernels<T>();
    const int n = (int)xitems.size();
    const unsigned grid = xcd_grid(std::min<long>(n, (long)devinfo().cus * 8));
    hipLaunchKernelGGL(wk.x, dim3(grid), dim3(wk.x_nt), wk.x_lds, stream, d_xitems, n, plans.d_plans, plans.d_tw,
                       st);
    HIPCHK(hipGetLastError());
  }
  void run(hipStream_t stream, const SimState* st) {
    const DevInfo& di = devinfo();
    run_planes(stream, st, false);
    for (int i = 0; i < 5; ++i) {
      if (i == 2) run_x(stream, st);
      const int n = offs[i + 1] - offs[i];
      if (n == 0) continue;
      require(lds[i] <= (size_t)di.lds, "FFT line too long for LDS: " + std::to_string(lds[i]) + " bytes");
      const unsigned grid = xcd_grid(std::min<long>(n, (long)di.cus * 32));
      if (nt == 64)
        hipLaunchKernelGGL((k_fft<T, 64>), dim3(grid), dim3(64), lds[i], stream, d_items + offs[i], n,
                           plans.d_plans, plans.d_tw, plans.d_roots, st);
      else if (nt == 512)
        hipLaunchKernelGGL((k_fft<T, 512>), dim3(grid), dim3(512), lds[i], stream, d_items + offs[i], n,
                           plans.d_plans, plans.d_tw, plans.d_roots, st);
      else if (nt == 256)
        hipLaunchKernelGGL((k_fft<T, 256>), dim3(grid), dim3(256), lds[i], stream, d_items + offs[i], n,
                           plans.d_plans, plans.d_tw, plans.d_roots, st);
      else
        hipLaunchKernelGGL((k_fft<T, 128>), dim3(grid), dim3(128), lds[i], stream, d_items + offs[i], n,
                           plans.d_plans, plans.d_tw, plans.d_roots, st);
      HIPCHK(hipGetLastError());
    }
    run_planes(stream, st, true);
  }
  ~FftBatch() {
    dfree(d_items);
    dfree(d_pitems);
    dfree(d_xitems);
  }
};

// Bhat (k_shape, double) -> device T in the spectrum layout, with 1/N folded in
template <typename T>
static T* upload_bhat(int dim, const long* shape, const double* bhat_host) {
  long rows = 1, ncell = 1;
  for (int a = 0; a < dim; ++a) ncell *= shape[a];
  for (int a = 0; a < dim - 1; ++a) rows *= shape[a];
  const long nh = shape[dim - 1] / 2 + 1, nhp = spec_row<T>(dim, shape);
  std::vector<T> h(rows * nhp, T(0));
  const double inv = 1.0 / (double)ncell;
  for (long r = 0; r < rows; ++r)
    for (long k = 0; k < nh; ++k) h[r * nhp + k] = (T)(bhat_host[r * nh + k] * inv);
  T* d = (T*)dmalloc(h.size() * sizeof(T));
  HIPCHK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return d;
}

// ---------------------------------------------------------------------------
// initial field generated on the device
// ---------------------------------------------------------------------------
// SplitMix64 (Steele, Lea and Flood, OOPSLA 2014) of a counter: uniform double
// in [0, 1) with 53 random bits, value number i of stream seed
PF5D_HD double unit_uniform(unsigned long long seed, unsigned long long i) {
  unsigned long long z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

constexpr int INIT_NT = 256;
// X = X0 + f (noise - mean(noise))[1:], noise = 2u - 1 over the five components
// (the distribution of apply_fluctuations, system.py:56-61, from a counter-based
// generator); FP32 mode stores the nearest FP32 value and the remainder.  Each
// block leaves Neumaier partial sums of the FP64 field (sum, compensation per
// component) in part[8 * block ...] for the FP64 mean.
template <typename T>
__global__ void __launch_bounds__(INIT_NT) k_init_field(T* __restrict__ X, Rem4* __restrict__ Xlo, long ncell,
                                                        double x00, double x01, double x02, double x03, double f,
                                                        unsigned long long seed, double* __restrict__ part) {
  const double X0[4] = {x00, x01, x02, x03};
  double sum[4] = {0, 0, 0, 0}, cmp[4] = {0, 0, 0, 0};
  for (long c = (long)blockIdx.x * INIT_NT + threadIdx.x; c < ncell; c += (long)gridDim.x * INIT_NT) {
    double n[5], m = 0.0;
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      n[k] = (unit_uniform(seed, (unsigned long long)c * 5 + k) - 0.5) * 2.0;
      m += n[k];
    }
    m /= 5.0;
    V4<T> hi;
    int lo[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const double x = X0[a] + f * (n[a + 1] - m);
      hi.v[a] = (T)x;
      lo[a] = rem_encode(x, (float)hi.v[a]);
      const double t = sum[a] + x;
      cmp[a] += (fabs(sum[a]) >= fabs(x)) ? ((sum[a] - t) + x) : ((x - t) + sum[a]);
      sum[a] = t;
    }
    st4g<T>(X + 4 * c, hi);
    if (Xlo) st_rem(Xlo + c, rem_pack(lo));
  }
  __shared__ double red[INIT_NT / 64][8];
  double v[8] = {sum[0], cmp[0], sum[1], cmp[1], sum[2], cmp[2], sum[3], cmp[3]};
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] += __shfl_xor(v[k], off, 64);
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 8; ++k) red[threadIdx.x >> 6][k] = v[k];
  __syncthreads();
  if (threadIdx.x < 8) {
    double r = 0.0;
    for (int w = 0; w < INIT_NT / 64; ++w) r += red[w][threadIdx.x];
    part[8 * blockIdx.x + threadIdx.x] = r;
  }
}

// ---------------------------------------------------------------------------
// plain elastic updater
// ---------------------------------------------------------------------------
struct pf5d_elastic {
  int prec, dim;
  long shape[3], ncell;
  double l[4], X0[4];
  void *s = nullptr, *C = nullptr, *psi = nullptr, *bhat = nullptr;
  std::unique_ptr<FftBatch<double>> b64;
  std::unique_ptr<FftBatch<float>> b32;
  ~pf5d_elastic() {
    dfree(s);
    dfree(C);
    dfree(psi);
    dfree(bhat);
  }
};

template <typename T>
static void elastic_init(pf5d_elastic* e, const double* bhat_host) {
  e->s = dmalloc(e->ncell * sizeof(T));
  e->psi = dmalloc(e->ncell * sizeof(T));
  e->C = dmalloc(spectrum_size<T>(e->dim, e->shape) * sizeof(cx<T>));
  e->bhat = upload_bhat<T>(e->dim, e->shape, bhat_host);
  auto b = std::make_unique<FftBatch<T>>();
  b->add_sim(0, e->dim, e->shape, (const T*)e->s, (cx<T>*)e->C, (T*)e->psi, (const T*)e->bhat);
  b->plans.upload();
  b->upload();
  if constexpr (sizeof(T) == 8)
    e->b64 = std::move(b);
  else
    e->b32 = std::move(b);
}

template <typename T>
static void elastic_apply(pf5d_elastic* e, void* mu, long mstride, const void* X, hipStream_t st) {
  V4<T> x0, l;
  for (int a = 0; a < 4; ++a) {
    x0.v[a] = (T)e->X0[a];
    l.v[a] = (T)e->l[a];
  }
  hipLaunchKernelGGL(k_misfit<T>, dim3(nblocks(e->ncell, 256)), dim3(256), 0, st, (const T*)X, e->ncell, x0, l,
                     (T*)e->s);
  HIPCHK(hipGetLastError());
  if constexpr (sizeof(T) == 8)
    e->b64->run(st, nullptr);
  else
    e->b32->run(st, nullptr);
  hipLaunchKernelGGL(k_add_psi<T>, dim3(nblocks(e->ncell, 256)), dim3(256), 0, st, (T*)mu, mstride,
                     (const T*)e->psi, e->ncell, l);
  HIPCHK(hipGetLastError());
}

// ---------------------------------------------------------------------------
// engine
// ---------------------------------------------------------------------------
struct EngineBase {
  virtual ~EngineBase() {}
  virtual void add(const pf5d_sim_desc* d, const double* X, const double* bhat, const pf5d_table* tab, int* slot) = 0;
  virtual void add_generated(const pf5d_sim_desc* d, unsigned long long seed, double fluctuation, const double* bhat,
                             const pf5d_table* tab, int* slot) = 0;
  virtual void remove(int slot) = 0;
  virtual void step(int n) = 0;
  virtual void prepare(int n) = 0;
  virtual void get_state(int slot, double* out) = 0;
  virtual int get_all_states(int maxn, double* out) = 0;
  virtual void set_time(int slot, double dt, double t) = 0;
  virtual void get_field(int slot, int which, double* out) = 0;
  virtual void set_field(int slot, const double* X) = 0;
  virtual void* field_ptr(int slot, int which) = 0;
  virtual void stats(long* active, long* bytes) = 0;
  virtual void set_tiling(int variant, int cx) = 0;
  virtual void set_fft(int budget_kb, int nt) = 0;
  virtual void set_fft_planes(int parts, int nt) = 0;
  virtual void set_fft_wave(int mode) = 0;
  virtual int num_variants(int* info) = 0;
  virtual void launch_info(long* out) = 0;
  virtual int num_groups() = 0;
  virtual double time_kernel(int which, int reps) = 0;
  virtual void time_step_stages(int reps, int separate_dt, double* out) = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
};

struct pf5d_engine {
  int prec;
  std::unique_ptr<EngineBase> impl;
};

template <typename T>
struct K1Variant {
  int by, bz, nt;  // potentials tiles of nt cells in rows of up to bz (by: 0)
  void (*fpot)(const SimDev<T>*, const SimState*, const K1Item*, int);
  void (*fflux)(const SimDev<T>*, const SimState*, T*, const K1Item*, int);
  // flux pass of the two-pass form: fcpl == 0 runs on the potentials pass's
  // per-plane tiles; fcpl > 0 marches along axis 0 over tiles of rows of
  // 64*fcpl cells, fnt/64 rows per block (fmarch, which can also run the dt
  // controller in its last block)
  int fcpl = 0, fnt = 0;
  void (*fmarch)(const SimDev<T>*, SimState*, T*, const K1Item*, int, const double*, const int*, int, const int*,
                 unsigned*) = nullptr;
  int fminw = 0;  // marching pass: min waves per SIMD it was compiled for
  // fused marching form (fmarch = k_stencilmarch): the potentials pass runs on
  // the boundary rows and planes of the marching tiles only, the marching pass
  // computes the other cells' mu and Mob itself
  bool fused = false;
  // potentials pass with the nearest-row search for tables with missing rows
  void (*fpot_wide)(const SimDev<T>*, const SimState*, const K1Item*, int) = nullptr;
};

// MINW: register budget of the potentials pass; LOADS: how the flux pass
// fetches the neighbour cells (0 per axis, 1 all up front, 2 up front with the
// z neighbours from adjacent lanes); BZ: row length of a tile
template <typename T, int NT, int MINW, int LOADS, int BZ = 64>
static K1Variant<T> make_split_variant() {
  K1Variant<T> v{0, BZ, NT, &k_potentials<T, NT, MINW, false>, &k_fluxdiv<T, NT, LOADS>, 0, NT};
  v.fpot_wide = &k_potentials<T, NT, MINW, true>;
  return v;
}
// two-pass form with the marching flux pass (FNT threads, CPL cells per lane,
// FMINW waves per SIMD at least)
// (the potentials pass's tiles are 16 x 16 cells: a compact tile keeps a
// developed field's cells close to its node box's anchor)
template <typename T, int NT, int MINW, int FNT, int CPL, int FMINW = 1, int BZ = 16>
static K1Variant<T> make_march_variant() {
  K1Variant<T> v{0, BZ, NT, &k_potentials<T, NT, MINW, false>, nullptr, CPL, FNT, &k_fluxmarch<T, FNT, CPL, FMINW>,
                 FMINW};
  v.fpot_wide = &k_potentials<T, NT, MINW, true>;
  return v;
}

// fused marching form: k_stencilmarch (FNT threads: FNT/64 rows per tile) with
// the potentials pass (NT threads, MINW waves per SIMD) on the tiles' boundary
// rows and planes
template <typename T, int NT, int MINW, int FNT, int FMINW>
static K1Variant<T> make_fused_variant() {
  K1Variant<T> v{0, 16, NT, &k_potentials<T, NT, MINW, false>, nullptr, 2, FNT, &k_stencilmarch<T, FNT, FMINW>,
                 FMINW};
  v.fpot_wide = &k_potentials<T, NT, MINW, true>;
  v.fused = true;
  return v;
}

// stencil tilings the autotuner chooses from
template <typename T>
static std::vector<K1Variant<T>> k1_variants();
template <>
std::vector<K1Variant<double>> k1_variants<double>() {
  return {make_split_variant<double, 256, 2, 0>(),
          make_split_variant<double, 256, 2, 1>(), make_split_variant<double, 256, 2, 1, 32>(),
          make_split_variant<double, 256, 2, 2>(), make_march_variant<double, 256, 2, 256, 1>(),
          make_march_variant<double, 256, 2, 256, 1, 3>()};
}
template <>
std::vector<K1Variant<float>> k1_variants<float>() {
  return {make_split_variant<float, 256, 3, 0>(), make_split_variant<float, 256, 4, 0>(),
          make_split_variant<float, 256, 4, 1>(), make_split_variant<float, 256, 4, 1, 32>(),
          make_split_variant<float, 256, 4, 2>(), make_march_variant<float, 256, 4, 256, 1>(),
          make_march_variant<float, 256, 4, 256, 2>(), make_march_variant<float, 256, 4, 512, 2>(),
          make_march_variant<float, 256, 4, 256, 2, 3>(), make_fused_variant<float, 256, 4, 512, 2>(),
          make_fused_variant<float, 256, 3, 512, 2>()};
}

constexpr int K2_NT = 256;
constexpr long K2_CHUNK = 2048;  // cells per K2 item: 1024 items (16 waves per CU) at 128^3
constexpr int DT_NT = 256;

template <typename T>
struct Engine : EngineBase {
  struct Sim {
    bool used = false;
    pf5d_sim_desc d{};
    long ncell = 0;
    T* X = nullptr;
    Rem4* Xlo = nullptr;
    T *dX = nullptr, *psi = nullptr, *s = nullptr, *bhat = nullptr;
    T *mu = nullptr, *mb = nullptr, *mb1 = nullptr;  // two-pass stencil intermediates
    cx<T>* C = nullptr;
    const pf5d_table* tab = nullptr;
    size_t bytes = 0;
    bool fresh = true;
    double mean0[4] = {0, 0, 0, 0};  // FP64 mean composition of the field as added
  };
  std::vector<Sim> sims;
  std::vector<K1Variant<T>> variants = k1_variants<T>();
  int variant = 0;
  int xchunk = 16;
  // x planes per potentials item of the marching two-pass form (3D): the item's
  // node box, tile set-up and barriers serve that many cells per thread.  At
  // 128^3, per-plane items take KP 37.5 / 69 us (FP32 / FP64); 2, 4 and 8 planes
  // take 31.8, 28.9, 27.1 us in FP32 and 62, 58, 63 us in FP64 (16 planes in
  // FP32, whose 512 items leave two blocks per CU: 33 us)
  int kp_planes = sizeof(T) == 4 ? 8 : 4;
  // device arrays
  SimDev<T>* d_sims = nullptr;
  SimState* d_state = nullptr;
  T* d_xmax = nullptr;
  unsigned* d_done = nullptr;  // block counter of the flux pass that runs the dt controller
  K1Item* d_k1 = nullptr;
  K1Item* d_kf = nullptr;  // tiles of the marching flux pass
  K2Item* d_k2 = nullptr;
  double* d_part = nullptr;
  int* d_prange = nullptr;
  int nk1 = 0, nkf = 0, nk2 = 0, cap = 0;
  FftBatch<T> fft;
  bool dirty = true;
  // Step graphs of 1, 2, 4, ... steps (graph gs[k]: 2^k steps).  A graph launch
  // leaves the GPU idle for about 8 us (128^3: 6 % of a step) while a graph runs
  // its kernels back to back, so n steps replay as many 2^GLOG-step graphs as
  // fit, then the binary digits of the rest (20 steps: 16 + 4).
  static constexpr int GLOG = 6;
  static constexpr int GLOGG = 4;  // groups: at most 16 steps per group graph (its fields stay cached meanwhile)
  hipGraphExec_t gs[GLOG + 1] = {};
  // the graph launches of n steps: gs[lmax] while 2^lmax steps are left, then the binary digits
  template <typename F>
  static void for_chunks(int n, int lmax, F f) {
    for (; n >= (1 << lmax); n -= 1 << lmax) f(lmax);
    for (int k = lmax - 1; k >= 0; --k)
      if (n & (1 << k)) f(k);
  }
  unsigned k1_grid = 0, kf_grid = 0;
  // Simulations whose per-step working sets together exceed what the 256 MB
  // Infinity Cache holds are stepped in groups: each group runs up to 16 steps (its
  // own graph) while its fields stay in the cache, then the next group.  Every
  // simulation advances by the same steps; only the order of independent work
  // changes.  A group is a range of the work lists, its own FFT batch and the
  // list of its simulations for the dt controller.  At most one group: the
  // whole engine steps as one graph.
  struct Group {
    std::vector<int> sims;
    int k1_off = 0, nk1 = 0, kf_off = 0, nkf = 0, k2_off = 0, nk2 = 0;
    unsigned k1_grid = 0, kf_grid = 0, k2_grid = 0;
    int* d_list = nullptr;
    FftBatch<T> fft;
    hipGraphExec_t gs[GLOGG + 1] = {};
    Group() = default;
    Group(const Group&) = delete;  // owns device memory and graphs
    Group& operator=(const Group&) = delete;
    void drop() {
      for (auto& g : gs) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
      }
    }
    ~Group() {
      drop();
      dfree(d_list);
    }
  };
  std::vector<std::unique_ptr<Group>> groups;
  // per-step working set, bytes per cell: X (+ 24-bit remainders), dX, mu, Mob, s,
  // psi and the spectrum; a group holds at most GROUP_BYTES of it
  static constexpr size_t STEP_BYTES = sizeof(T) == 4 ? 96 : 170;
  static constexpr size_t GROUP_BYTES = size_t(192) << 20;

  Engine(hipStream_t s) {
    if (s) {
      stream = s;
    } else {
      HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
      own_stream = true;
    }
    d_done = (unsigned*)dmalloc(sizeof(unsigned));
    HIPCHK(hipMemset(d_done, 0, sizeof(unsigned)));
    // default tiling (until set_tiling): the marching two-pass form with 16 x 16
    // potentials tiles, the flux pass with two cells per lane at two waves per
    // SIMD in FP32 and one cell at three in FP64, the autotuner's pick for 3D
    // fields from 64^3 to 256^3; x chunks of 16 planes; FFT lines of about six points per thread
    // (32 KiB, 256 threads) and two (FP32) or four (FP64) parts per plane of the
    // 128 x 128 plane passes
    const int want_cpl = sizeof(T) == 4 ? 2 : 1, want_minw = sizeof(T) == 4 ? 1 : 3;
    variant = -1;
    for (int i = 0; i < (int)variants.size() && variant < 0; ++i)
      if (variants[i].fcpl == want_cpl && variants[i].fminw == want_minw && variants[i].fnt == 256 &&
          variants[i].bz == 16)
        variant = i;
    require(variant >= 0, "no default stencil tiling");
    default_variant = variant;
    fft.budget = 32 * 1024;
    fft.nt = 256;
    fft.parts = sizeof(T) == 4 ? 2 : 4;
    fft.pnt = 1024;
    fft.wave = sizeof(T) == 4 ? 3 : FFT_WAVE_ALL;  // (FP32: the Stockham inverse plane pass is faster)
    (void)hipFuncSetAttribute((const void*)&k_fft<T, 64>, hipFuncAttributeMaxDynamicSharedMemorySize, devinfo().lds);
    (void)hipFuncSetAttribute((const void*)&k_fft<T, 512>, hipFuncAttributeMaxDynamicSharedMemorySize, devinfo().lds);
    (void)hipFuncSetAttribute((const void*)&k_fft<T, 256>, hipFuncAttributeMaxDynamicSharedMemorySize, devinfo().lds);
    (void)hipFuncSetAttribute((const void*)&k_fft<T, 128>, hipFuncAttributeMaxDynamicSharedMemorySize, devinfo().lds);
    for (int pn : {64, 96, 128})
      for (int pnt : {256, 512, 1024})
        for (int parts : {2, 4, 8}) {
          const PlaneKernels<T> k = plane_kernels<T>(pn, parts, pnt);
          if (!k.fwd) continue;
          (void)hipFuncSetAttribute((const void*)k.fwd, hipFuncAttributeMaxDynamicSharedMemorySize, devinfo().lds);
          (void)hipFuncSetAttribute((const void*)k.inv, hipFuncAttributeMaxDynamicSharedMemorySize, devinfo().lds);
        }
    {
      const WaveKernels<T> k = wave_kernels<T>();
      (void)hipFuncSetAttribute((const void*)k.fwd, hipFuncAttributeMaxDynamicSharedMemorySize, devinfo().lds);
      (void)hipFuncSetAttribute((const void*)k.inv, hipFuncAttributeMaxDynamicSharedMemorySize, devinfo().lds);
      (void)hipFuncSetAttribute((const void*)k.x, hipFuncAttributeMaxDynamicSharedMemorySize, devinfo().lds);
    }
  }
  ~Engine() override {
    drop_graphs();
    for (auto& s : sims) free_sim(s);
    dfree(d_done);
    dfree(d_sims);
    dfree(d_state);
    dfree(d_xmax);
    dfree(d_k1);
    dfree(d_kf);
    dfree(d_k2);
    dfree(d_part);
    dfree(d_prange);
    if (own_stream) (void)hipStreamDestroy(stream);
  }
  void free_sim(Sim& s) {
    dfree(s.X);
    dfree(s.Xlo);
    dfree(s.dX);
    dfree(s.psi);
    dfree(s.s);
    dfree(s.bhat);
    dfree(s.C);
    dfree(s.mu);
    dfree(s.mb);
    dfree(s.mb1);
    s = Sim();
  }
  void drop_graphs() {
    for (auto& g : gs) {
      if (g) (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
    for (auto& g : groups) g->drop();
  }
  void ensure_capacity(int n) {
    if (n <= cap) return;
    int nc = std::max(n, std::max(8, 2 * cap));
    SimState* ns = (SimState*)dmalloc(nc * sizeof(SimState));
    HIPCHK(hipMemset(ns, 0, nc * sizeof(SimState)));
    if (d_state) HIPCHK(hipMemcpy(ns, d_state, cap * sizeof(SimState), hipMemcpyDeviceToDevice));
    dfree(d_state);
    d_state = ns;
    T* nx = (T*)dmalloc(nc * sizeof(T));
    HIPCHK(hipMemset(nx, 0, nc * sizeof(T)));
    dfree(d_xmax);
    d_xmax = nx;
    dfree(d_sims);
    d_sims = (SimDev<T>*)dmalloc(nc * sizeof(SimDev<T>));
    dfree(d_prange);
    d_prange = (int*)dmalloc(nc * 2 * sizeof(int));
    cap = nc;
  }
  void add(const pf5d_sim_desc* d, const double* Xh, const double* bhat, const pf5d_table* tab, int* slot) override {
    add_with(d, bhat, tab, slot, [&](Sim& s) { upload_X(s, Xh); });
  }
  void add_generated(const pf5d_sim_desc* d, unsigned long long seed, double fluctuation, const double* bhat,
                     const pf5d_table* tab, int* slot) override {
    add_with(d, bhat, tab, slot, [&](Sim& s) {
      const unsigned grid = (unsigned)std::max<long>(1, std::min<long>((s.ncell + INIT_NT - 1) / INIT_NT, 4096));
      double* part = (double*)dmalloc((size_t)grid * 8 * sizeof(double));
      try {
        hipLaunchKernelGGL(k_init_field<T>, dim3(grid), dim3(INIT_NT), 0, stream, s.X, s.Xlo, s.ncell, d->X0[0],
                           d->X0[1], d->X0[2], d->X0[3], fluctuation, seed, part);
        HIPCHK(hipGetLastError());
        std::vector<double> h((size_t)grid * 8);
        HIPCHK(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        for (int a = 0; a < 4; ++a) {
          double sum = 0, c = 0;
          for (unsigned b = 0; b < grid; ++b)
            for (int q = 0; q < 2; ++q) {
              const double v = h[8 * b + 2 * a + q], t = sum + v;
              c += (std::fabs(sum) >= std::fabs(v)) ? ((sum - t) + v) : ((v - t) + sum);
              sum = t;
            }
          s.mean0[a] = (sum + c) / (double)s.ncell;
        }
      } catch (...) {
        dfree(part);
        throw;
      }
      dfree(part);
    });
  }
  template <typename Fill>
  void add_with(const pf5d_sim_desc* d, const double* bhat, const pf5d_table* tab, int* slot, Fill fill) {
    require(d->dim >= 1 && d->dim <= 3, "dim must be 1, 2 or 3", PF5D_ENOTIMPL);
    require(tab && tab->prec == (sizeof(T) == 8 ? PF5D_F64 : PF5D_F32), "table precision does not match engine");
    require(tab->num >= BOX4, "the engine needs a table with at least 4 nodes per axis");
    long ncell = 1;
    for (int a = 0; a < d->dim; ++a) {
      require(d->shape[a] >= 4, "every axis needs at least 4 cells");
      ncell *= d->shape[a];
    }
    // the stencil kernels address a simulation's (ncell, 4) fields with 32-bit
    // byte offsets: up to 2^28 cells in FP32 (645^3), 2^27 in FP64 (512^3)
    require(ncell * 4 * (long)sizeof(T) <= (1L << 32),
            "field too large: at most " + std::to_string((1L << 32) / (4 * (long)sizeof(T))) + " cells per simulation");
    require(!d->elastic || bhat, "elastic simulation needs bhat");
    HIPCHK(hipStreamSynchronize(stream));
    int id = -1;
    for (int i = 0; i < (int)sims.size(); ++i)
      if (!sims[i].used) {
        id = i;
        break;
      }
    if (id < 0) {
      id = (int)sims.size();
      sims.emplace_back();
    }
    ensure_capacity((int)sims.size());
    Sim& s = sims[id];
    s.used = true;
    s.d = *d;
    s.ncell = ncell;
    s.tab = tab;
    s.fresh = true;
    try {
      s.X = (T*)dmalloc(ncell * 4 * sizeof(T));
      s.dX = (T*)dmalloc(ncell * 4 * sizeof(T));
      s.mu = (T*)dmalloc(ncell * 4 * sizeof(T));
      s.mb = (T*)dmalloc(ncell * 4 * sizeof(T));
      s.mb1 = (T*)dmalloc(ncell * sizeof(T));
      s.bytes = ncell * 17 * sizeof(T);
      if (sizeof(T) == 4) {
        s.Xlo = (Rem4*)dmalloc(ncell * sizeof(Rem4));
        s.bytes += ncell * sizeof(Rem4);
      }
      if (d->elastic) {
        s.psi = (T*)dmalloc(ncell * sizeof(T));
        s.s = (T*)dmalloc(ncell * sizeof(T));
        s.C = (cx<T>*)dmalloc(spectrum_size<T>(d->dim, d->shape) * sizeof(cx<T>));
        s.bhat = upload_bhat<T>(d->dim, d->shape, bhat);
        s.bytes += ncell * 2 * sizeof(T) + spectrum_size<T>(d->dim, d->shape) * (sizeof(cx<T>) + sizeof(T));
      }
      fill(s);
      HIPCHK(hipMemset(s.dX, 0, ncell * 4 * sizeof(T)));
      // psi is written by the first step's FFT passes; zeroed so that nothing
      // reads uninitialised memory before (a stencil timed right after add)
      if (s.psi) HIPCHK(hipMemset(s.psi, 0, ncell * sizeof(T)));
    } catch (...) {
      free_sim(s);
      throw;
    }
    SimState st{};
    st.dt = d->dt0;
    st.t = 0.0;
    st.status = ST_ACTIVE;
    HIPCHK(hipMemcpy(d_state + id, &st, sizeof(SimState), hipMemcpyHostToDevice));
    *slot = id;
    dirty = true;
  }
  // FP32 mode keeps X as the nearest FP32 value plus a 24-bit remainder
  void upload_X(Sim& s, const double* Xh) {
    // FP64 mean of the field (pairwise over cells is unnecessary at this size: Neumaier)
    for (int a = 0; a < 4; ++a) {
      double sum = 0, c = 0;
      for (long i = 0; i < s.ncell; ++i) {
        const double v = Xh[4 * i + a], t = sum + v;
        c += (std::fabs(sum) >= std::fabs(v)) ? ((sum - t) + v) : ((v - t) + sum);
        sum = t;
      }
      s.mean0[a] = (sum + c) / (double)s.ncell;
    }
    std::vector<T> h(s.ncell * 4);
    std::vector<Rem4> lo(s.Xlo ? s.ncell : 0);
    for (long c = 0; c < s.ncell; ++c) {
      int v[4];
      for (int a = 0; a < 4; ++a) {
        h[4 * c + a] = (T)Xh[4 * c + a];
        v[a] = rem_encode(Xh[4 * c + a], (float)h[4 * c + a]);
      }
      if (s.Xlo) lo[c] = rem_pack(v);
    }
    HIPCHK(hipMemcpy(s.X, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    if (s.Xlo) HIPCHK(hipMemcpy(s.Xlo, lo.data(), lo.size() * sizeof(Rem4), hipMemcpyHostToDevice));
  }
  Sim& at(int slot) {
    require(slot >= 0 && slot < (int)sims.size() && sims[slot].used, "invalid slot " + std::to_string(slot));
    return sims[slot];
  }
  void remove(int slot) override {
    Sim& s = at(slot);
    HIPCHK(hipStreamSynchronize(stream));
    free_sim(s);
    SimState st{};
    HIPCHK(hipMemcpy(d_state + slot, &st, sizeof(SimState), hipMemcpyHostToDevice));
    dirty = true;
  }
  SimDev<T> make_dev(const Sim& s) {
    SimDev<T> v{};
    const pf5d_sim_desc& d = s.d;
    v.dim = d.dim;
    if (d.dim == 3) {
      v.n0 = (int)d.shape[0];
      v.n1 = (int)d.shape[1];
      v.n2 = (int)d.shape[2];
    } else if (d.dim == 2) {
      v.n0 = 1;
      v.n1 = (int)d.shape[0];
      v.n2 = (int)d.shape[1];
    } else {
      v.n0 = 1;
      v.n1 = 1;
      v.n2 = (int)d.shape[0];
    }
    v.ncell = s.ncell;
    v.inv_dx = (T)(1.0 / d.cell_size);
    v.inv_dx2 = (T)(1.0 / (d.cell_size * d.cell_size));
    for (int a = 0; a < 4; ++a) {
      v.kappa[a] = (T)d.kappa[a];
      v.l[a] = d.elastic ? (T)d.l[a] : T(0);
      // FP32 mode: misfit about the conserved FP64 mean, so the FFT input carries no k=0 mode
      v.X0[a] = (sizeof(T) == 4) ? (T)s.mean0[a] : (T)d.X0[a];
    }
    v.elastic = d.elastic;
    v.use_mean = (sizeof(T) == 4) ? 1 : 0;
    v.fresh = s.fresh ? 1 : 0;
    v.tab = view<T>(s.tab);
    v.X = s.X;
    v.Xlo = s.Xlo;
    v.dX = s.dX;
    v.psi = s.psi;
    v.s = s.s;
    v.mu = s.mu;
    v.mb = s.mb;
    v.mb1 = s.mb1;
    v.ctld = d.ctld;
    v.dt_cfl = d.dt_cfl;
    v.lower = d.lower_limit;
    v.upper = d.upper_limit;
    v.t_end = d.t_end;
    v.oob_interval = d.oob_interval;
    v.tol_soft = d.tol_soft;
    v.tol_hard = d.tol_hard;
    return v;
  }
  // a simulation's table misses rows of the simplex: its nearest-row search must
  // skip them (the WIDE potentials pass)
  bool wide = false;
  int default_variant = 0;
  void (*pot(const K1Variant<T>& V) const)(const SimDev<T>*, const SimState*, const K1Item*, int) {
    return wide ? V.fpot_wide : V.fpot;
  }
  // The fused marching form (k_stencilmarch) needs every simulation 3D with
  // rows of 64 * fcpl cells, the other axes in whole tiles (fnt / 64 rows) and
  // runs (xchunk >= 3 planes), and complete tables (the wide nearest-row search
  // is in the potentials pass only).
  bool fused_fits(const K1Variant<T>& V) const {
    if (!V.fused || wide || xchunk < 3) return false;
    const int rows = V.fnt / 64;
    int n = 0;
    for (const auto& s : sims) {
      if (!s.used) continue;
      ++n;
      const pf5d_sim_desc& d = s.d;
      if (d.dim != 3 || d.shape[2] != 64 * V.fcpl || d.shape[1] % rows != 0 || d.shape[0] % xchunk != 0)
        return false;
    }
    return n > 0;
  }
  // Work lists of the fused marching form for simulation i: marching tiles of
  // `rows` rows over runs of R planes (kf), and potentials items (k1) on the
  // cells whose mu and Mob the marching pass reads but does not compute -- the
  // runs' first and last planes (every row), and on the other planes the tiles'
  // first and last rows.  Both lists in slabs of whole runs, one per XCD, in the
  // order the other kernels' lists take them.
  void fused_lists(int i, const SimDev<T>& v, const K1Variant<T>& V, std::vector<K1Item>& k1,
                   std::vector<K1Item>& kf) const {
    const int R = xchunk, rows = V.fnt / 64, n0 = v.n0, n1 = v.n1, n2 = v.n2;
    const int nruns = n0 / R, nsl = std::min(NXCD, nruns);
    const int tz = std::min(V.bz, n2), ry = V.nt / tz;  // potentials tiles of the boundary planes
    for (int sl = 0; sl < nsl; ++sl) {
      const int ka = sl * nruns / nsl, kb = (sl + 1) * nruns / nsl;
      for (int k = ka; k < kb; ++k) {
        const int xa = k * R;
        // the plane below the run and its first plane (wrapping at x = 0)
        auto planes = [&](int pa, int pb) {
          for (int z0 = 0; z0 < n2; z0 += tz)
            for (int y0 = 0; y0 < n1; y0 += ry) k1.push_back({i, pa, pb, y0, z0, std::min(ry, n1 - y0), tz});
        };
        if (xa > 0) {
          planes(xa - 1, xa + 1);
        } else {
          planes(n0 - 1, n0);
          planes(0, 1);
        }
        // the tiles' last and first rows on the run's inner planes: row pairs
        // (rows*j - 1, rows*j), rows of n2 cells, in items of about the cells
        // of a plane item (the persistent grid's blocks then take similar work)
        const int per = std::max(1, 2 * (tz * ry) / (2 * n2));  // planes per row item
        for (int x0 = xa + 1; x0 < xa + R - 1; x0 += per) {
          const int x1 = std::min(x0 + per, xa + R - 1);
          for (int y = rows - 1; y < n1; y += rows) {
            if (y + 1 < n1) {
              k1.push_back({i, x0, x1, y, 0, 2, n2});
            } else {
              k1.push_back({i, x0, x1, y, 0, 1, n2});
              k1.push_back({i, x0, x1, 0, 0, 1, n2});
            }
          }
        }
      }
      for (int y0 = 0; y0 < n1; y0 += rows)
        for (int k = ka; k < kb; ++k) kf.push_back({i, k * R, k * R + R, y0, 0, rows, n2});
    }
  }
  void rebuild() {
    HIPCHK(hipStreamSynchronize(stream));  // buffers below may still be in use
    drop_graphs();
    wide = false;
    for (const auto& s : sims)
      if (s.used && s.tab && !s.tab->complete) wide = true;
    // the fused marching form takes only the shapes it was built for
    if (variants[variant].fused && !fused_fits(variants[variant])) variant = default_variant;
    const K1Variant<T>& V = variants[variant];
    std::vector<SimDev<T>> devs(cap);
    std::vector<K1Item> k1, kf;
    std::vector<K2Item> k2;
    std::vector<int> prange(2 * cap, 0);
    fft.clear();
    groups.clear();
    // groups (marching two-pass form only): consecutive simulations up to
    // GROUP_BYTES of working set, at least one each
    std::vector<std::vector<int>> gsims;
    {
      const bool can_group = V.fcpl > 0;
      size_t acc = 0;
      for (int i = 0; i < (int)sims.size(); ++i) {
        if (!sims[i].used) continue;
        const size_t b = (size_t)sims[i].ncell * STEP_BYTES;
        if (gsims.empty() || (can_group && !gsims.back().empty() && acc + b > GROUP_BYTES)) {
          gsims.emplace_back();
          acc = 0;
        }
        gsims.back().push_back(i);
        acc += b;
      }
    }
    std::vector<int> order, gstart;  // simulations in group order; first of each group in order
    for (auto& g : gsims) {
      gstart.push_back((int)order.size());
      order.insert(order.end(), g.begin(), g.end());
    }
    std::vector<int> k1_at(order.size() + 1), kf_at(order.size() + 1), k2_at(order.size() + 1);
    for (int oi = 0; oi < (int)order.size(); ++oi) {
      const int i = order[oi];
      Sim& s = sims[i];
      k1_at[oi] = (int)k1.size();
      kf_at[oi] = (int)kf.size();
      k2_at[oi] = (int)k2.size();
      devs[i] = make_dev(s);
      const SimDev<T>& v = devs[i];
      const bool hx = v.dim == 3;
      if (V.fused) {
        fused_lists(i, v, V, k1, kf);
      } else {
        // tiles of at most nt cells: rows of up to bz cells (nt in 1D), as many rows as fit
        // slabs of planes (one per XCD, in the order the other kernels' lists
        // take them), then tile columns, x fastest: a block's contiguous range
        // of items walks down one tile column of one slab
        const int tzmax = (v.n1 == 1) ? V.nt : V.bz;
        // planes per potentials item (sharing one staged node box); the per-plane
        // flux pass (fcpl == 0) walks the same items one plane each
        const int kpx = (V.fcpl > 0 && hx) ? kp_planes : 1;
        const int nsl = std::min(NXCD, v.n0);
        for (int sl = 0; sl < nsl; ++sl) {
          const int xa = sl * v.n0 / nsl, xb = (sl + 1) * v.n0 / nsl;
          for (int z0 = 0; z0 < v.n2; z0 += tzmax) {
            const int tz = std::min(tzmax, v.n2 - z0), ry = V.nt / tz;
            for (int y0 = 0; y0 < v.n1; y0 += ry)
              for (int x = xa; x < xb; x += kpx)
                k1.push_back({i, x, std::min(x + kpx, xb), y0, z0, std::min(ry, v.n1 - y0), tz});
          }
        }
        if (V.fcpl > 0) {
          // marching flux tiles: fnt/64 rows of up to 64*fcpl cells, runs of
          // xchunk planes, in the same slab order
          const int rz = 64 * V.fcpl, ry = V.fnt / 64;
          const int cxs = hx ? std::max(1, xchunk) : 1;
          for (int sl = 0; sl < nsl; ++sl) {
            const int xa = sl * v.n0 / nsl, xb = (sl + 1) * v.n0 / nsl;
            for (int z0 = 0; z0 < v.n2; z0 += rz)
              for (int y0 = 0; y0 < v.n1; y0 += ry)
                for (int x = xa; x < xb; x += cxs)
                  kf.push_back({i, x, std::min(x + cxs, xb), y0, z0, std::min(ry, v.n1 - y0), std::min(rz, v.n2 - z0)});
          }
        }
      }
      prange[2 * i] = (int)k2.size();
      for (long c0 = 0; c0 < s.ncell; c0 += K2_CHUNK) k2.push_back({i, c0, std::min(s.ncell, c0 + K2_CHUNK)});
      prange[2 * i + 1] = (int)k2.size();
      if (s.d.elastic)
        fft.add_sim(i, s.d.dim, s.d.shape, s.s, s.C, s.psi, s.bhat);
    }
    k1_at[order.size()] = (int)k1.size();
    kf_at[order.size()] = (int)kf.size();
    k2_at[order.size()] = (int)k2.size();
    fft.plans.upload();
    fft.upload();
    if (gsims.size() > 1) {
      for (size_t gi = 0; gi < gsims.size(); ++gi) {
        auto g = std::make_unique<Group>();
        g->sims = gsims[gi];
        const int a = gstart[gi], b = a + (int)g->sims.size();
        g->k1_off = k1_at[a];
        g->nk1 = k1_at[b] - k1_at[a];
        g->kf_off = kf_at[a];
        g->nkf = kf_at[b] - kf_at[a];
        g->k2_off = k2_at[a];
        g->nk2 = k2_at[b] - k2_at[a];
        g->fft.budget = fft.budget;
        g->fft.nt = fft.nt;
        g->fft.parts = fft.parts;
        g->fft.pnt = fft.pnt;
        g->fft.wave = fft.wave;
        for (int i : g->sims) {
          const Sim& s = sims[i];
          if (s.d.elastic) g->fft.add_sim(i, s.d.dim, s.d.shape, s.s, s.C, s.psi, s.bhat);
        }
        g->fft.plans.upload();
        g->fft.upload();
        g->d_list = (int*)dmalloc(g->sims.size() * sizeof(int));
        HIPCHK(hipMemcpy(g->d_list, g->sims.data(), g->sims.size() * sizeof(int), hipMemcpyHostToDevice));
        groups.push_back(std::move(g));
      }
    }
    HIPCHK(hipMemcpy(d_sims, devs.data(), cap * sizeof(SimDev<T>), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(d_prange, prange.data(), prange.size() * sizeof(int), hipMemcpyHostToDevice));
    dfree(d_k1);
    dfree(d_kf);
    dfree(d_k2);
    dfree(d_part);
    d_k1 = nullptr;
    d_kf = nullptr;
    d_k2 = nullptr;
    d_part = nullptr;
    nk1 = (int)k1.size();
    nkf = (int)kf.size();
    nk2 = (int)k2.size();
    if (nkf) {
      d_kf = (K1Item*)dmalloc(nkf * sizeof(K1Item));
      HIPCHK(hipMemcpy(d_kf, kf.data(), nkf * sizeof(K1Item), hipMemcpyHostToDevice));
    }
    if (nk1) {
      d_k1 = (K1Item*)dmalloc(nk1 * sizeof(K1Item));
      HIPCHK(hipMemcpy(d_k1, k1.data(), nk1 * sizeof(K1Item), hipMemcpyHostToDevice));
    }
    if (nk2) {
      d_k2 = (K2Item*)dmalloc(nk2 * sizeof(K2Item));
      HIPCHK(hipMemcpy(d_k2, k2.data(), nk2 * sizeof(K2Item), hipMemcpyHostToDevice));
      d_part = (double*)dmalloc((size_t)nk2 * NSTAT * sizeof(double));
    }
    // K1 grid: persistent, at most what fits on the chip at once
    int occ = 1;
    {
      auto grid = [&](int o) { return xcd_grid(std::min(nk1, devinfo().cus * std::max(1, o))); };
      HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)pot(V), V.nt, 0));
      // (the fused form's potentials items are small and alike: one block each,
      // and the dispatcher keeps the CUs evenly loaded)
      k1_grid = V.fused ? xcd_grid(nk1) : grid(occ);
      occ = 1;
      const void* ff = V.fcpl > 0 ? (const void*)V.fmarch : (const void*)V.fflux;
      HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, ff, V.fnt, 0));
      kf_grid = V.fcpl > 0 ? xcd_grid(std::min(nkf, devinfo().cus * std::max(1, occ))) : grid(occ);
      int occ_pot = 1;
      HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ_pot, (const void*)pot(V), V.nt, 0));
      for (auto& g : groups) {
        g->k1_grid = V.fused ? xcd_grid(g->nk1)
                             : xcd_grid(std::max(1, std::min(g->nk1, devinfo().cus * std::max(1, occ_pot))));
        g->kf_grid = xcd_grid(std::max(1, std::min(g->nkf, devinfo().cus * std::max(1, occ))));
        g->k2_grid = xcd_grid(std::max(1, std::min(g->nk2, devinfo().cus * 8)));
      }
    }
    // prime after every rebuild (the partial-sum buffer is new): FP64 statistics
    // of the current X, and the misfit of fields just set; X and the other
    // misfits are not modified, so a rebuild between two steps leaves every
    // simulation's trajectory unchanged
    if (nk2) {
      launch_k2(1);
      launch_dt(0);
    }
    // later prime launches (state reads) keep the misfit the steps write
    bool any_fresh = false;
    for (auto& v : devs) {
      any_fresh = any_fresh || v.fresh;
      v.fresh = 0;
    }
    if (any_fresh) {
      HIPCHK(hipMemcpyAsync(d_sims, devs.data(), cap * sizeof(SimDev<T>), hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
    for (auto& s : sims) s.fresh = false;
    dirty = false;
    stats_stale = false;
  }
  void launch_k2(int prime) {
    if (!nk2) return;
    const unsigned grid = xcd_grid(std::min(nk2, devinfo().cus * 8));
    hipLaunchKernelGGL((k_update<T, K2_NT>), dim3(grid), dim3(K2_NT), 0, stream, d_sims, d_state, d_k2, nk2, d_part,
                       prime);
    HIPCHK(hipGetLastError());
  }
  void launch_dt(int advance) {
    hipLaunchKernelGGL((k_dtctl<T, DT_NT>), dim3((unsigned)sims.size()), dim3(DT_NT), 0, stream, d_sims, d_state, d_xmax,
                       d_part, d_prange, (int)sims.size(), advance, nullptr);
    HIPCHK(hipGetLastError());
  }
  // the flux pass; with_dt: the marching pass also runs the dt controller
  // (returns whether it did)
  bool launch_flux(const K1Variant<T>& V, bool with_dt = false) {
    if (V.fcpl > 0) {
      const int ndt = (with_dt && fuse_dt_ok()) ? (int)sims.size() : 0;
      hipLaunchKernelGGL(V.fmarch, dim3(kf_grid), dim3(V.fnt), 0, stream, d_sims, d_state, d_xmax, d_kf, nkf, d_part,
                         d_prange, ndt, nullptr, d_done);
      return ndt > 0;
    }
    hipLaunchKernelGGL(V.fflux, dim3(kf_grid), dim3(V.nt), 0, stream, d_sims, d_state, d_xmax, d_k1, nk1);
    return false;
  }
  // the flux pass's last block runs the controllers of a few simulations (a
  // launch of one block per simulation is better for many)
  bool fuse_dt_ok() const { return sims.size() <= 8 && nk2 > 0; }
  // K1; with_dt: the dt controller too where the flux pass can run it (returns whether it did)
  bool launch_k1(bool with_dt = false) {
    if (!nk1) return false;
    const K1Variant<T>& V = variants[variant];
    hipLaunchKernelGGL(pot(V), dim3(k1_grid), dim3(V.nt), 0, stream, d_sims, d_state, d_k1, nk1);
    HIPCHK(hipGetLastError());
    const bool did = launch_flux(V, with_dt);
    HIPCHK(hipGetLastError());
    return did;
  }
  void enqueue_step() {
    fft.run(stream, d_state);
    if (!launch_k1(true)) launch_dt(1);
    launch_k2(0);
  }
  // one step of one group (marching two-pass form): its FFT passes, potentials,
  // flux with the dt controller of its simulations, update
  void enqueue_group_step(Group& g) {
    const K1Variant<T>& V = variants[variant];
    g.fft.run(stream, d_state);
    if (g.nk1) {
      hipLaunchKernelGGL(pot(V), dim3(g.k1_grid), dim3(V.nt), 0, stream, d_sims, d_state, d_k1 + g.k1_off, g.nk1);
      HIPCHK(hipGetLastError());
    }
    const int ns = (int)g.sims.size();
    const int ndt = (g.nkf && g.nk2 && ns <= 8) ? ns : 0;
    if (g.nkf) {
      hipLaunchKernelGGL(V.fmarch, dim3(g.kf_grid), dim3(V.fnt), 0, stream, d_sims, d_state, d_xmax, d_kf + g.kf_off,
                         g.nkf, d_part, d_prange, ndt, g.d_list, d_done);
      HIPCHK(hipGetLastError());
    }
    if (!ndt && g.nk2) {
      hipLaunchKernelGGL((k_dtctl<T, DT_NT>), dim3((unsigned)ns), dim3(DT_NT), 0, stream, d_sims, d_state, d_xmax,
                         d_part, d_prange, ns, 1, g.d_list);
      HIPCHK(hipGetLastError());
    }
    if (g.nk2) {
      hipLaunchKernelGGL((k_update<T, K2_NT>), dim3(g.k2_grid), dim3(K2_NT), 0, stream, d_sims, d_state,
                         d_k2 + g.k2_off, g.nk2, d_part + (long)g.k2_off * NSTAT, 0);
      HIPCHK(hipGetLastError());
    }
  }
  template <typename F>
  hipGraphExec_t capture_with(int n, F enqueue) {
    hipGraph_t g;
    HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    try {
      for (int i = 0; i < n; ++i) enqueue();
    } catch (...) {
      (void)hipStreamEndCapture(stream, &g);
      throw;
    }
    HIPCHK(hipStreamEndCapture(stream, &g));
    hipGraphExec_t ex;
    HIPCHK(hipGraphInstantiate(&ex, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    return ex;
  }
  hipGraphExec_t capture(int n) {
    return capture_with(n, [&] { enqueue_step(); });
  }
  int num_groups() override {
    if (dirty) rebuild();
    return std::max<int>(1, (int)groups.size());
  }
  // the step graphs step(n) replays, captured (not run) ahead of it
  void prepare(int n) override {
    require(n >= 0, "nsteps must be >= 0");
    if (dirty) rebuild();
    if (nk1 == 0 || n == 0) return;
    if (groups.size() > 1) {
      for (auto& gp : groups) {
        Group& g = *gp;
        for_chunks(n, GLOGG, [&](int k) {
          if (!g.gs[k]) g.gs[k] = capture_with(1 << k, [&] { enqueue_group_step(g); });
        });
      }
      return;
    }
    for_chunks(n, GLOG, [&](int k) {
      if (!gs[k]) gs[k] = capture(1 << k);
    });
  }
  void step(int n) override {
    prepare(n);
    if (nk1 == 0 || n == 0) return;
    if (groups.size() > 1) {
      // each chunk: every group in turn (a group's fields stay cached through its chunk)
      for_chunks(n, GLOGG, [&](int k) {
        for (auto& g : groups) HIPCHK(hipGraphLaunch(g->gs[k], stream));
      });
      stats_stale = true;
      return;
    }
    for_chunks(n, GLOG, [&](int k) { HIPCHK(hipGraphLaunch(gs[k], stream)); });
    stats_stale = true;
  }
  // steps compute the FP64 statistics only ahead of bounds checks: bring them
  // up to date (statistics of the current fields, nothing advanced) before a
  // state is read
  bool stats_stale = false;
  void refresh_stats() {
    if (!stats_stale) return;
    if (nk2) {
      launch_k2(1);
      launch_dt(0);
    }
    stats_stale = false;
  }
  void get_state(int slot, double* out) override {
    at(slot);
    if (dirty) rebuild();
    refresh_stats();
    SimState st;
    HIPCHK(hipMemcpyAsync(&st, d_state + slot, sizeof(SimState), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    const double v[16] = {st.dt,  st.t,    st.last_xmax, st.mass[0], st.mass[1], st.mass[2],     st.mass[3], (double)st.steps,
                          (double)st.status, st.smax, st.smin, st.cmin, st.cmax, st.nonfinite, st.mean[0], st.mean[1]};
    std::memcpy(out, v, sizeof(v));
  }
  // all slots in one copy: out[16*i ...] for slot i < nslots (unused slots: status 0)
  int get_all_states(int maxn, double* out) override {
    if (dirty) rebuild();
    refresh_stats();
    const int n = std::min(maxn, (int)sims.size());
    if (n <= 0) return 0;
    std::vector<SimState> h(n);
    HIPCHK(hipMemcpyAsync(h.data(), d_state, n * sizeof(SimState), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    for (int i = 0; i < n; ++i) {
      const SimState& st = h[i];
      const double v[16] = {st.dt,          st.t,    st.last_xmax, st.mass[0], st.mass[1], st.mass[2],
                            st.mass[3],     (double)st.steps,      (double)(sims[i].used ? st.status : 0),
                            st.smax,        st.smin, st.cmin,      st.cmax,    st.nonfinite, st.mean[0], st.mean[1]};
      std::memcpy(out + 16 * i, v, sizeof(v));
    }
    return n;
  }
  void set_time(int slot, double dt, double t) override {
    at(slot);
    if (dirty) rebuild();
    HIPCHK(hipStreamSynchronize(stream));
    SimState st;
    HIPCHK(hipMemcpy(&st, d_state + slot, sizeof(SimState), hipMemcpyDeviceToHost));
    st.dt = dt;
    st.t = t;
    HIPCHK(hipMemcpy(d_state + slot, &st, sizeof(SimState), hipMemcpyHostToDevice));
  }
  void get_field(int slot, int which, double* out) override {
    Sim& s = at(slot);
    const T* src = which == 0 ? s.X : which == 1 ? s.dX : s.psi;
    require(src != nullptr, "field not present");
    const long n = (which == 2) ? s.ncell : s.ncell * 4;
    std::vector<T> h(n);
    std::vector<Rem4> lo((which == 0 && s.Xlo) ? s.ncell : 0);
    HIPCHK(hipMemcpyAsync(h.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, stream));
    if (!lo.empty()) HIPCHK(hipMemcpyAsync(lo.data(), s.Xlo, lo.size() * sizeof(Rem4), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (lo.empty()) {
      for (long i = 0; i < n; ++i) out[i] = (double)h[i];
    } else {
      for (long c = 0; c < s.ncell; ++c) {
        int v[4];
        rem_unpack(lo[c], v);
        for (int a = 0; a < 4; ++a) out[4 * c + a] = rem_decode((float)h[4 * c + a], v[a]);
      }
    }
  }
  void set_field(int slot, const double* X) override {
    Sim& s = at(slot);
    HIPCHK(hipStreamSynchronize(stream));
    upload_X(s, X);
    s.fresh = true;
    dirty = true;
  }
  void* field_ptr(int slot, int which) override {
    Sim& s = at(slot);
    return which == 0 ? (void*)s.X : which == 1 ? (void*)s.dX : (void*)s.psi;
  }
  void stats(long* active, long* bytes) override {
    long a = 0, b = 0;
    for (auto& s : sims)
      if (s.used) {
        ++a;
        b += (long)s.bytes;
      }
    *active = a;
    *bytes = b;
  }
  void set_tiling(int v, int c) override {
    require(v >= 0 && v < (int)variants.size(), "tiling variant out of range");
    require(c >= 1, "cx must be >= 1");
    variant = v;
    xchunk = c;
    dirty = true;
  }
  void set_fft_planes(int parts, int nt) override {
    require(parts == 0 || parts == 2 || parts == 4 || parts == 8, "FFT plane parts must be 0, 2, 4 or 8");
    require(nt == 256 || nt == 512 || nt == 1024, "FFT plane block size must be 256, 512 or 1024");
    fft.parts = parts;
    fft.pnt = nt;
    dirty = true;
  }
  void set_fft_wave(int mode) override {
    require(mode >= 0 && mode <= FFT_WAVE_ALL, "FFT wave mode must be a mask in 0.." + std::to_string(FFT_WAVE_ALL));
    fft.wave = mode;
    dirty = true;
  }
  void set_fft(int budget_kb, int nt) override {
    require(budget_kb >= 4 && budget_kb <= 160, "FFT LDS budget must be 4..160 KiB");
    require(nt == 64 || nt == 128 || nt == 256 || nt == 512, "FFT block size must be 64, 128, 256 or 512");
    fft.budget = (size_t)budget_kb * 1024;
    fft.nt = nt;
    dirty = true;
  }
  int num_variants(int* info) override {
    if (info)
      for (size_t i = 0; i < variants.size(); ++i) {
        info[5 * i] = variants[i].fused ? -1 : variants[i].by;
        info[5 * i + 1] = variants[i].bz;
        info[5 * i + 2] = variants[i].nt;
        info[5 * i + 3] = variants[i].fcpl;
        info[5 * i + 4] = variants[i].fnt;
      }
    return (int)variants.size();
  }
  // average duration of one launch of a single stage, timed with HIP events on
  // the engine stream: 0 = K1 stencil, 1 = FFT stages (all five), 2 = DT, 3 = K2 prime,
  // 4 / 5 = potentials / flux pass of the two-pass stencil; 6 = the FFT passes, K1
  // and a K2 prime in sequence (the stencil stage between the passes a step runs
  // around it, whose traffic leaves the caches as a step does)
  double time_kernel(int which, int reps) override {
    require(reps > 0 && which >= 0 && which <= 6, "time_kernel: bad arguments");
    if (dirty) rebuild();
    const K1Variant<T>& V = variants[variant];
    hipEvent_t a, b;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    auto one = [&] {
      if (which == 0) launch_k1();
      else if (which == 1) fft.run(stream, d_state);
      else if (which == 2) launch_dt(0);
      else if (which == 3) launch_k2(1);
      else if (which == 4)
        hipLaunchKernelGGL(pot(V), dim3(k1_grid), dim3(V.nt), 0, stream, d_sims, d_state, d_k1, nk1);
      else if (which == 5)
        launch_flux(V);
      else {
        fft.run(stream, d_state);
        launch_k1();
        launch_k2(1);
      }
    };
    one();
    HIPCHK(hipEventRecord(a, stream));
    for (int i = 0; i < reps; ++i) one();
    HIPCHK(hipEventRecord(b, stream));
    // the timed stencil folded max|dX/dt| of the current fields into the
    // accumulator the next step's dt controller reads, and a timed controller
    // reduced whatever partials the update pass last left: restore both (the
    // accumulator to zero, the statistics to those of the current fields)
    if (nk2) {
      launch_k2(1);
      launch_dt(0);
      stats_stale = false;
    }
    HIPCHK(hipEventSynchronize(b));
    HIPCHK(hipStreamSynchronize(stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return (double)ms / reps;
  }
  // per-stage device time of whole steps: reps steps launched one by one (not
  // from the step graph) on the engine stream, with HIP events between the
  // stages; out[0..5] = average ms of the FFT passes, potentials pass, flux
  // pass (with the dt controller in its last block unless separate_dt),
  // separate dt launch (0 when fused), update pass, whole step.  The steps
  // advance the simulations exactly as step(reps) does.
  void time_step_stages(int reps, int separate_dt, double* out) override {
    require(reps > 0 && reps <= 4096, "time_step_stages: reps must be in 1..4096");
    if (dirty) rebuild();
    require(groups.size() <= 1, "time_step_stages: the simulations form more than one group");
    const K1Variant<T>& V = variants[variant];
    require(V.fcpl > 0, "time_step_stages: the marching two-pass stencil is not selected");
    require(nk1 > 0 && nk2 > 0, "time_step_stages: no active simulation");
    constexpr int NE = 6;
    std::vector<hipEvent_t> ev((size_t)NE * reps, nullptr);
    auto destroy = [&] {
      for (auto& x : ev)
        if (x) (void)hipEventDestroy(x);
    };
    try {
      for (auto& x : ev) HIPCHK(hipEventCreate(&x));
      bool fused = false;
      for (int r = 0; r < reps; ++r) {
        hipEvent_t* e = &ev[(size_t)NE * r];
        HIPCHK(hipEventRecord(e[0], stream));
        fft.run(stream, d_state);
        HIPCHK(hipEventRecord(e[1], stream));
        hipLaunchKernelGGL(pot(V), dim3(k1_grid), dim3(V.nt), 0, stream, d_sims, d_state, d_k1, nk1);
        HIPCHK(hipGetLastError());
        HIPCHK(hipEventRecord(e[2], stream));
        fused = launch_flux(V, !separate_dt);
        HIPCHK(hipGetLastError());
        if (!fused) {
          HIPCHK(hipEventRecord(e[3], stream));
          launch_dt(1);
        }
        HIPCHK(hipEventRecord(e[4], stream));
        launch_k2(0);
        HIPCHK(hipEventRecord(e[5], stream));
      }
      stats_stale = true;
      HIPCHK(hipEventSynchronize(ev.back()));
      double acc[NE] = {0, 0, 0, 0, 0, 0};
      for (int r = 0; r < reps; ++r) {
        hipEvent_t* e = &ev[(size_t)NE * r];
        float ms = 0;
        for (int i = 0; i + 1 < NE; ++i) {
          if (fused && i == 3) continue;  // no separate dt launch: the flux stage ends at e[4]
          HIPCHK(hipEventElapsedTime(&ms, e[i], e[(fused && i == 2) ? 4 : i + 1]));
          acc[i] += ms;
        }
        HIPCHK(hipEventElapsedTime(&ms, e[0], e[NE - 1]));
        acc[NE - 1] += ms;
      }
      for (int i = 0; i < NE; ++i) out[i] = acc[i] / reps;
    } catch (...) {
      destroy();
      throw;
    }
    destroy();
  }
  void launch_info(long* out) override {
    if (dirty) rebuild();
    out[0] = nk1;
    out[1] = nk2;
    out[2] = fft.offs[5];
    out[3] = k1_grid;
    out[4] = 0;  // (K1 LDS bytes: every stencil pass has static LDS only)
    out[5] = variant;
    out[6] = xchunk;
    out[7] = (long)fft.budget;
    out[8] = fft.nt;
    out[9] = kf_grid;
    out[10] = fft.parts_used;
    out[11] = fft.pnt;
    out[12] = fft.wave_used;
  }
};

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
static hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }
static void check_prec(int p) { require(p == PF5D_F64 || p == PF5D_F32, "prec must be PF5D_F64 or PF5D_F32"); }
static Shape make_shape(int dim, const long* shape) {
  if (dim < 1 || dim > 3) throw Pf5dError(PF5D_ENOTIMPL, "dimension must be 1, 2 or 3");
  Shape s;
  s.dim = dim;
  s.ncell = 1;
  for (int a = 0; a < 3; ++a) s.n[a] = a < dim ? (int)shape[a] : 1;
  for (int a = 0; a < dim; ++a) {
    require(shape[a] >= 1, "shape entries must be positive");
    s.ncell *= shape[a];
  }
  return s;
}

extern "C" {

const char* pf5d_last_error(void) { return g_err.c_str(); }
int pf5d_version(void) { return 1; }

int pf5d_device_info(int out[3]) {
  return guarded([&] {
    const DevInfo& d = devinfo();
    out[0] = d.cus;
    out[1] = d.lds;
    out[2] = d.warp;
  });
}

int pf5d_table_create(int prec, int num, int nprops, const double* grid, double resolution, pf5d_table** out) {
  return guarded([&] {
    check_prec(prec);
    require(num >= 2 && nprops == 9 && grid && out && resolution > 0, "table: need num >= 2, nprops == 9");
    auto t = std::make_unique<pf5d_table>();
    t->prec = prec;
    t->num = num;
    t->resolution = resolution;
    t->data = nullptr;
    t->cubes = nullptr;
    t->cube_s = -1;
    if (prec == PF5D_F64)
      upload_table<double>(t.get(), grid, nprops);
    else
      upload_table<float>(t.get(), grid, nprops);
    *out = t.release();
  });
}
int pf5d_table_destroy(pf5d_table* t) {
  return guarded([&] {
    if (!t) return;
    dfree(t->data);
    dfree(t->cubes);
    delete t;
  });
}

int pf5d_interpolate(int prec, const void* X, long ncell, const pf5d_table* tab, void* out, long ostride,
                     void* stream) {
  return guarded([&] {
    check_prec(prec);
    require(tab && tab->prec == prec, "table precision mismatch");
    require(ostride >= 9, "output row stride must be >= 9");
    if (ncell <= 0) return;
    if (prec == PF5D_F64)
      hipLaunchKernelGGL(k_interp<double>, dim3(nblocks(ncell, 256)), dim3(256), 0, S(stream), (const double*)X,
                         ncell, view<double>(tab), (double*)out, ostride);
    else
      hipLaunchKernelGGL(k_interp<float>, dim3(nblocks(ncell, 256)), dim3(256), 0, S(stream), (const float*)X, ncell,
                         view<float>(tab), (float*)out, ostride);
    HIPCHK(hipGetLastError());
  });
}

int pf5d_laplacian(int prec, const void* X, int dim, const long* shape, double h, void* out, void* stream) {
  return guarded([&] {
    check_prec(prec);
    Shape sh = make_shape(dim, shape);
    const double inv = 1.0 / (h * h);
    if (prec == PF5D_F64)
      hipLaunchKernelGGL(k_laplacian<double>, dim3(nblocks(sh.ncell, 256)), dim3(256), 0, S(stream),
                         (const double*)X, sh, inv, (double*)out);
    else
      hipLaunchKernelGGL(k_laplacian<float>, dim3(nblocks(sh.ncell, 256)), dim3(256), 0, S(stream), (const float*)X,
                         sh, (float)inv, (float*)out);
    HIPCHK(hipGetLastError());
  });
}

int pf5d_kappa_subtract(int prec, void* L, void* mu, long mstride, long ncell, const double kappa[4], void* stream) {
  return guarded([&] {
    check_prec(prec);
    if (ncell <= 0) return;
    if (prec == PF5D_F64) {
      V4<double> k{{kappa[0], kappa[1], kappa[2], kappa[3]}};
      hipLaunchKernelGGL(k_kappa_sub<double>, dim3(nblocks(ncell, 256)), dim3(256), 0, S(stream), (double*)L,
                         (double*)mu, mstride, ncell, k);
    } else {
      V4<float> k{{(float)kappa[0], (float)kappa[1], (float)kappa[2], (float)kappa[3]}};
      hipLaunchKernelGGL(k_kappa_sub<float>, dim3(nblocks(ncell, 256)), dim3(256), 0, S(stream), (float*)L,
                         (float*)mu, mstride, ncell, k);
    }
    HIPCHK(hipGetLastError());
  });
}

int pf5d_divergence_from_mu(int prec, const void* X, const void* mob, long bstride, const void* mu, long mstride,
                            int dim, const long* shape, double inv_dx, void* out, void* stream) {
  return guarded([&] {
    check_prec(prec);
    Shape sh = make_shape(dim, shape);
    if (prec == PF5D_F64)
      hipLaunchKernelGGL(k_div_from_mu<double>, dim3(nblocks(sh.ncell, 256)), dim3(256), 0, S(stream),
                         (const double*)X, (const double*)mob, bstride, (const double*)mu, mstride, sh, inv_dx,
                         (double*)out);
    else
      hipLaunchKernelGGL(k_div_from_mu<float>, dim3(nblocks(sh.ncell, 256)), dim3(256), 0, S(stream),
                         (const float*)X, (const float*)mob, bstride, (const float*)mu, mstride, sh, (float)inv_dx,
                         (float*)out);
    HIPCHK(hipGetLastError());
  });
}

int pf5d_gradients_pm(int prec, const void* A, long astride, int dim, const long* shape, double h, void* grad_p,
                      void* grad_m, void* stream) {
  return guarded([&] {
    check_prec(prec);
    require(astride >= 4 && h > 0, "gradients: need row stride >= 4 and h > 0");
    Shape sh = make_shape(dim, shape);
    if (prec == PF5D_F64)
      hipLaunchKernelGGL(k_gradients_pm<double>, dim3(nblocks(sh.ncell, 256)), dim3(256), 0, S(stream),
                         (const double*)A, astride, sh, 1.0 / h, (double*)grad_p, (double*)grad_m);
    else
      hipLaunchKernelGGL(k_gradients_pm<float>, dim3(nblocks(sh.ncell, 256)), dim3(256), 0, S(stream),
                         (const float*)A, astride, sh, (float)(1.0 / h), (float*)grad_p, (float*)grad_m);
    HIPCHK(hipGetLastError());
  });
}

int pf5d_fluxes_pm(int prec, const void* X, const void* mob, long bstride, int dim, const long* shape,
                   const void* grad_p, const void* grad_m, void* fluxes_p, void* fluxes_m, void* stream) {
  return guarded([&] {
    check_prec(prec);
    require(bstride >= 5, "fluxes: mobility row stride must be >= 5");
    Shape sh = make_shape(dim, shape);
    if (prec == PF5D_F64)
      hipLaunchKernelGGL(k_fluxes_pm<double>, dim3(nblocks(sh.ncell, 256)), dim3(256), 0, S(stream),
                         (const double*)X, (const double*)mob, bstride, sh, (const double*)grad_p,
                         (const double*)grad_m, (double*)fluxes_p, (double*)fluxes_m);
    else
      hipLaunchKernelGGL(k_fluxes_pm<float>, dim3(nblocks(sh.ncell, 256)), dim3(256), 0, S(stream),
                         (const float*)X, (const float*)mob, bstride, sh, (const float*)grad_p,
                         (const float*)grad_m, (float*)fluxes_p, (float*)fluxes_m);
    HIPCHK(hipGetLastError());
  });
}

int pf5d_field_stats(int prec, const void* A, long ncell, long astride, int ncomp, double* out, void* stream) {
  return guarded([&] {
    check_prec(prec);
    require(ncell > 0 && ncomp >= 1 && ncomp <= 8 && astride >= ncomp && out, "field_stats: bad arguments");
    const unsigned grid = (unsigned)std::min<long>(nblocks(ncell, 256), 4L * devinfo().cus);
    double* d = (double*)grow_scratch().get((size_t)(64 + grid * 24) * sizeof(double));
    double *acc = d, *mean = d + 40, *part = d + 64, *qpart = part + (size_t)grid * 16;
    hipStream_t st = S(stream);
    hipLaunchKernelGGL(k_stats_init, dim3(1), dim3(64), 0, st, acc);
    if (prec == PF5D_F64)
      hipLaunchKernelGGL(k_stats<double>, dim3(grid), dim3(256), 0, st, (const double*)A, ncell, astride, ncomp,
                         part, acc);
    else
      hipLaunchKernelGGL(k_stats<float>, dim3(grid), dim3(256), 0, st, (const float*)A, ncell, astride, ncomp,
                         part, acc);
    HIPCHK(hipGetLastError());
    std::vector<double> hp((size_t)grid * 16), ha(33);
    HIPCHK(hipMemcpyAsync(hp.data(), part, hp.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(ha.data(), acc, ha.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    double m[8];
    for (int k = 0; k < ncomp; ++k) {
      double sum = 0, c = 0;
      for (unsigned b = 0; b < grid; ++b) {
        for (double v : {hp[b * 16 + k], hp[b * 16 + 8 + k]}) {
          const double t = sum + v;
          c += (std::fabs(sum) >= std::fabs(v)) ? ((sum - t) + v) : ((v - t) + sum);
          sum = t;
        }
      }
      out[k] = sum + c;
      m[k] = out[k] / (double)ncell;
      out[ncomp + k] = ha[16 + k];
      out[2 * ncomp + k] = ha[24 + k];
    }
    out[4 * ncomp] = ha[32];
    HIPCHK(hipMemcpyAsync(mean, m, ncomp * sizeof(double), hipMemcpyHostToDevice, st));
    if (prec == PF5D_F64)
      hipLaunchKernelGGL(k_sqdev<double>, dim3(grid), dim3(256), 0, st, (const double*)A, ncell, astride, ncomp,
                         mean, qpart);
    else
      hipLaunchKernelGGL(k_sqdev<float>, dim3(grid), dim3(256), 0, st, (const float*)A, ncell, astride, ncomp,
                         mean, qpart);
    HIPCHK(hipGetLastError());
    std::vector<double> hq((size_t)grid * 8);
    HIPCHK(hipMemcpyAsync(hq.data(), qpart, hq.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int k = 0; k < ncomp; ++k) {
      double q = 0;
      for (unsigned b = 0; b < grid; ++b) q += hq[b * 8 + k];
      out[3 * ncomp + k] = q / (double)ncell;
    }
  });
}

int pf5d_field_dot(int prec, const void* A, const void* B, long ncell, int ncomp, double* out, void* stream) {
  return guarded([&] {
    check_prec(prec);
    require(ncell > 0 && ncomp >= 1 && ncomp <= 8 && out, "field_dot: bad arguments");
    const unsigned grid = (unsigned)std::min<long>(nblocks(ncell, 256), 4L * devinfo().cus);
    double* part = (double*)grow_scratch().get((size_t)grid * 8 * sizeof(double));
    hipStream_t st = S(stream);
    if (prec == PF5D_F64)
      hipLaunchKernelGGL(k_dot<double>, dim3(grid), dim3(256), 0, st, (const double*)A, (const double*)B, ncell,
                         ncomp, part);
    else
      hipLaunchKernelGGL(k_dot<float>, dim3(grid), dim3(256), 0, st, (const float*)A, (const float*)B, ncell, ncomp,
                         part);
    HIPCHK(hipGetLastError());
    std::vector<double> h((size_t)grid * 8);
    HIPCHK(hipMemcpyAsync(h.data(), part, h.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int k = 0; k < ncomp; ++k) {
      double q = 0;
      for (unsigned b = 0; b < grid; ++b) q += h[b * 8 + k];
      out[k] = q;
    }
  });
}

int pf5d_abs_max(int prec, const void* a, long n, double* host_result, void* stream) {
  return guarded([&] {
    check_prec(prec);
    void* d = grow_scratch().get(16);
    const unsigned grid = (unsigned)std::min<long>(nblocks(n, 256), 4L * devinfo().cus);
    double r = 0;
    if (prec == PF5D_F64) {
      hipLaunchKernelGGL(k_fill<double>, dim3(1), dim3(64), 0, S(stream), (double*)d, 1L, 0.0);
      if (n > 0)
        hipLaunchKernelGGL(k_absmax<double>, dim3(grid), dim3(256), 0, S(stream), (const double*)a, n, (double*)d);
      HIPCHK(hipMemcpyAsync(&r, d, sizeof(double), hipMemcpyDeviceToHost, S(stream)));
      HIPCHK(hipStreamSynchronize(S(stream)));
    } else {
      float f = 0;
      hipLaunchKernelGGL(k_fill<float>, dim3(1), dim3(64), 0, S(stream), (float*)d, 1L, 0.0f);
      if (n > 0)
        hipLaunchKernelGGL(k_absmax<float>, dim3(grid), dim3(256), 0, S(stream), (const float*)a, n, (float*)d);
      HIPCHK(hipMemcpyAsync(&f, d, sizeof(float), hipMemcpyDeviceToHost, S(stream)));
      HIPCHK(hipStreamSynchronize(S(stream)));
      r = f;
    }
    HIPCHK(hipGetLastError());
    *host_result = r;
  });
}

int pf5d_update(int prec, void* X, void* dX, long n, double dt, void* stream) {
  return guarded([&] {
    check_prec(prec);
    if (n <= 0) return;
    if (prec == PF5D_F64)
      hipLaunchKernelGGL(k_update_plain<double>, dim3(nblocks(n, 256)), dim3(256), 0, S(stream), (double*)X,
                         (double*)dX, n, dt);
    else
      hipLaunchKernelGGL(k_update_plain<float>, dim3(nblocks(n, 256)), dim3(256), 0, S(stream), (float*)X,
                         (float*)dX, n, dt);
    HIPCHK(hipGetLastError());
  });
}

int pf5d_clamp_simplex(int prec, void* X, long ncell, void* stream) {
  return guarded([&] {
    check_prec(prec);
    if (ncell <= 0) return;
    if (prec == PF5D_F64)
      hipLaunchKernelGGL(k_clamp<double>, dim3(nblocks(ncell, 256)), dim3(256), 0, S(stream), (double*)X, ncell);
    else
      hipLaunchKernelGGL(k_clamp<float>, dim3(nblocks(ncell, 256)), dim3(256), 0, S(stream), (float*)X, ncell);
    HIPCHK(hipGetLastError());
  });
}

}  // extern "C"

// Blocks of the bounds statistics: one per 256 cells, at most two per CU.
static unsigned bounds_grid(long ncell) { return (unsigned)std::min<long>(nblocks(ncell, 256), 2L * devinfo().cus); }

// Bounds statistics of X with the per-block partials in device buffer d (grid * 5
// doubles) and their host copy in h (as many): every block writes its own five
// partials, so nothing needs clearing; the host folds them in block order.
static void bounds_stats_into(int prec, const void* X, long ncell, double out5[5], hipStream_t st, double* d,
                              double* h, unsigned grid) {
  if (prec == PF5D_F64)
    hipLaunchKernelGGL(k_oob_partial<double>, dim3(grid), dim3(256), 0, st, (const double*)X, ncell, d);
  else
    hipLaunchKernelGGL(k_oob_partial<float>, dim3(grid), dim3(256), 0, st, (const float*)X, ncell, d);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(h, d, (size_t)grid * 5 * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  double v[5] = {0, -1e300, 1e300, 1e300, -1e300};
  for (unsigned b = 0; b < grid; ++b) {
    v[0] += h[b * 5];
    v[1] = std::max(v[1], h[b * 5 + 1]);
    v[2] = std::min(v[2], h[b * 5 + 2]);
    v[3] = std::min(v[3], h[b * 5 + 3]);
    v[4] = std::max(v[4], h[b * 5 + 4]);
  }
  std::memcpy(out5, v, sizeof(v));
}

// Fixed device scratch of pf5d_bounds_stats_noalloc: the per-block partials of
// up to NOALLOC_MAX_BLOCKS blocks (40 KiB), one copy per GPU for the life of the
// process.  Host code never passes `scratch` itself to a kernel (in host code it
// names a host-side stand-in); it uses the device address that
// hipGetSymbolAddress returns for the current device, kept in scratch_dev.
constexpr unsigned NOALLOC_MAX_BLOCKS = 1024;
static __device__ __align__(256) double scratch[NOALLOC_MAX_BLOCKS * 5];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_dev[NOALLOC_MAX_DEVICES] = {};

// device address of `scratch` on the current device (looked up once per device)
static double* scratch_dev() {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  require(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "bounds_stats_noalloc: device index out of range", PF5D_EDEVICE);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw Pf5dError(PF5D_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<double*>(p);
  }
  return g_scratch_dev[dev];
}

extern "C" {

int pf5d_bounds_stats(int prec, const void* X, long ncell, double out5[5], void* stream) {
  return guarded([&] {
    check_prec(prec);
    const unsigned grid = bounds_grid(ncell);
    double* d = (double*)grow_scratch().get(grid * 5 * sizeof(double));
    std::vector<double> h(grid * 5);
    bounds_stats_into(prec, X, ncell, out5, S(stream), d, h.data(), grid);
  });
}

int pf5d_bounds_stats_noalloc(int prec, const void* X, long ncell, double out5[5], void* stream) {
  return guarded([&] {
    check_prec(prec);
    require(X && out5, "bounds_stats_noalloc: null argument");
    const unsigned grid = bounds_grid(ncell);
    if (grid > NOALLOC_MAX_BLOCKS)
      throw Pf5dError(PF5D_EINVAL, "bounds_stats_noalloc: " + std::to_string(grid) + " blocks of partials exceed the " +
                                       std::to_string(NOALLOC_MAX_BLOCKS) + " the fixed scratch holds");
    double h[NOALLOC_MAX_BLOCKS * 5];
    bounds_stats_into(prec, X, ncell, out5, S(stream), scratch_dev(), h, grid);
  });
}

int pf5d_elastic_create(int prec, int dim, const long* shape, const double* bhat, const double l[4],
                        const double X0[4], pf5d_elastic** out) {
  return guarded([&] {
    check_prec(prec);
    Shape sh = make_shape(dim, shape);
    require(bhat && l && X0 && out, "elastic: null argument");
    auto e = std::make_unique<pf5d_elastic>();
    e->prec = prec;
    e->dim = dim;
    for (int a = 0; a < 3; ++a) e->shape[a] = a < dim ? shape[a] : 1;
    e->ncell = sh.ncell;
    for (int a = 0; a < 4; ++a) {
      e->l[a] = l[a];
      e->X0[a] = X0[a];
    }
    if (prec == PF5D_F64)
      elastic_init<double>(e.get(), bhat);
    else
      elastic_init<float>(e.get(), bhat);
    *out = e.release();
  });
}
int pf5d_elastic_apply(pf5d_elastic* e, void* mu, long mstride, const void* X, void* stream) {
  return guarded([&] {
    require(e != nullptr, "null elastic handle");
    if (e->prec == PF5D_F64)
      elastic_apply<double>(e, mu, mstride, X, S(stream));
    else
      elastic_apply<float>(e, mu, mstride, X, S(stream));
  });
}
const void* pf5d_elastic_psi(const pf5d_elastic* e) { return e ? e->psi : nullptr; }
int pf5d_elastic_destroy(pf5d_elastic* e) {
  return guarded([&] { delete e; });
}

int pf5d_engine_create(int prec, void* stream, pf5d_engine** out) {
  return guarded([&] {
    check_prec(prec);
    require(out != nullptr, "null output");
    auto e = std::make_unique<pf5d_engine>();
    e->prec = prec;
    if (prec == PF5D_F64)
      e->impl = std::make_unique<Engine<double>>(S(stream));
    else
      e->impl = std::make_unique<Engine<float>>(S(stream));
    *out = e.release();
  });
}
int pf5d_engine_destroy(pf5d_engine* e) {
  return guarded([&] {
    if (e && e->impl) (void)hipStreamSynchronize(e->impl->stream);
    delete e;
  });
}
int pf5d_engine_add(pf5d_engine* e, const pf5d_sim_desc* d, const double* X, const double* bhat,
                    const pf5d_table* tab, int* slot) {
  return guarded([&] {
    require(e && d && X && slot, "engine_add: null argument");
    e->impl->add(d, X, bhat, tab, slot);
  });
}
int pf5d_engine_add_generated(pf5d_engine* e, const pf5d_sim_desc* d, unsigned long long seed, double fluctuation,
                              const double* bhat, const pf5d_table* tab, int* slot) {
  return guarded([&] {
    require(e && d && slot, "engine_add_generated: null argument");
    e->impl->add_generated(d, seed, fluctuation, bhat, tab, slot);
  });
}
int pf5d_engine_remove(pf5d_engine* e, int slot) {
  return guarded([&] { e->impl->remove(slot); });
}
int pf5d_engine_num_variants(pf5d_engine* e, int* n, int* info) {
  return guarded([&] { *n = e->impl->num_variants(info); });
}
int pf5d_engine_set_fft_planes(pf5d_engine* e, int parts, int nt) {
  return guarded([&] { e->impl->set_fft_planes(parts, nt); });
}
int pf5d_engine_set_fft_wave(pf5d_engine* e, int mode) {
  return guarded([&] { e->impl->set_fft_wave(mode); });
}
int pf5d_engine_set_fft(pf5d_engine* e, int budget_kb, int nt) {
  return guarded([&] { e->impl->set_fft(budget_kb, nt); });
}
int pf5d_engine_set_tiling(pf5d_engine* e, int variant, int cx) {
  return guarded([&] { e->impl->set_tiling(variant, cx); });
}
int pf5d_engine_prepare(pf5d_engine* e, int nsteps) {
  return guarded([&] { e->impl->prepare(nsteps); });
}
int pf5d_engine_step(pf5d_engine* e, int nsteps) {
  return guarded([&] { e->impl->step(nsteps); });
}
int pf5d_engine_sync(pf5d_engine* e) {
  return guarded([&] { HIPCHK(hipStreamSynchronize(e->impl->stream)); });
}
int pf5d_engine_get_state(pf5d_engine* e, int slot, double* st) {
  return guarded([&] { e->impl->get_state(slot, st); });
}
int pf5d_engine_get_all_states(pf5d_engine* e, int maxn, double* out, int* n) {
  return guarded([&] { *n = e->impl->get_all_states(maxn, out); });
}
int pf5d_engine_set_time(pf5d_engine* e, int slot, double dt, double t) {
  return guarded([&] { e->impl->set_time(slot, dt, t); });
}
int pf5d_engine_get_field(pf5d_engine* e, int slot, int which, double* out) {
  return guarded([&] {
    require(which >= 0 && which <= 2, "which must be 0, 1 or 2");
    e->impl->get_field(slot, which, out);
  });
}
int pf5d_engine_set_field(pf5d_engine* e, int slot, const double* X) {
  return guarded([&] { e->impl->set_field(slot, X); });
}
void* pf5d_engine_field_ptr(pf5d_engine* e, int slot, int which) {
  void* p = nullptr;
  int rc = guarded([&] { p = e->impl->field_ptr(slot, which); });
  return rc == PF5D_OK ? p : nullptr;
}
void* pf5d_engine_stream(pf5d_engine* e) { return e ? (void*)e->impl->stream : nullptr; }
int pf5d_engine_stats(pf5d_engine* e, long* active, long* bytes) {
  return guarded([&] { e->impl->stats(active, bytes); });
}
int pf5d_engine_time_kernel(pf5d_engine* e, int which, int reps, double* avg_ms) {
  return guarded([&] { *avg_ms = e->impl->time_kernel(which, reps); });
}
int pf5d_engine_time_step_stages(pf5d_engine* e, int reps, int separate_dt, double* out) {
  return guarded([&] { e->impl->time_step_stages(reps, separate_dt, out); });
}
int pf5d_engine_num_groups(pf5d_engine* e, int* ngroups) {
  return guarded([&] { *ngroups = e->impl->num_groups(); });
}

int pf5d_engine_launch_info(pf5d_engine* e, long out[13]) {
  return guarded([&] { e->impl->launch_info(out); });
}

}  // extern "C"
