// Face flux with the dyadic face mobility, and the per-function ("plain")
// kernels that mirror the reference's Numba functions one to one.
#pragma once
#include "pf5d_common.h"
#include "pf5d_interp.h"

namespace pf5d {

// Flux through the face between cells a and b for potential gradient g
// (phasefield5d/solver/fluxes.py:128-160).  Xm = (Xa+Xb)/2, Mm = (Ma+Mb)/2,
// w_Fe = (1 - sum Xm) Mm_Fe, d = Xm.g,
// J_a = w_Fe (-Xm_a)(-d) + Xm_a Mm_a (g_a - d) - Xm_a sum_k Xm_k Mm_k (g_k - d),
// i.e. M_face g with M = w_Fe X X^T + sum_k Xm_k Mm_k (e_k - X)(e_k - X)^T.
// The expression is symmetric in (a, b), so the flux a face contributes to
// its two cells is the same value, which makes the divergence telescope.
template <typename T>
PF5D_DEV void face_flux(const T* Xa, const T* Xb, const T* Ma, const T* Mb, const T* g, T* out) {
  T Xm[4], Mm[5];
#pragma unroll
  for (int a = 0; a < 4; ++a) Xm[a] = T(0.5) * (Xa[a] + Xb[a]);
#pragma unroll
  for (int m = 0; m < 5; ++m) Mm[m] = T(0.5) * (Ma[m] + Mb[m]);
  const T sumX = ((Xm[0] + Xm[1]) + Xm[2]) + Xm[3];
  const T w_fe = (T(1) - sumX) * Mm[0];
  T dot = Xm[0] * g[0];
  dot = mad<T>(Xm[1], g[1], dot);
  dot = mad<T>(Xm[2], g[2], dot);
  dot = mad<T>(Xm[3], g[3], dot);
  const T dot_fe = -dot;
  T dk[4], wk[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    dk[k] = g[k] - dot;
    wk[k] = Xm[k] * Mm[1 + k];
  }
  T swk = wk[0] * dk[0];
  swk = mad<T>(wk[1], dk[1], swk);
  swk = mad<T>(wk[2], dk[2], swk);
  swk = mad<T>(wk[3], dk[3], swk);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    T v = (w_fe * (-Xm[a])) * dot_fe;
    v = mad<T>(wk[a], dk[a], v);
    out[a] = mad<T>(-Xm[a], swk, v);
  }
}

// The step kernels' form of the face flux.  FP64 is face_flux above with
// g = (mu_b - mu_a) / h, in the reference's operation order.  FP32 (a
// precision form with a tolerance, not the reference's bits) drops the two
// constant factors -- g = mu_b - mu_a, face mobility Ma + Mb instead of half of
// it -- since the flux is linear in both, and regroups the algebra as
//   J_a = Xm_a (w_Fe d - S) + Xm_a Mm_a (g_a - d),   S = sum_k Xm_k Mm_k (g_k - d),
// 47 operations per face instead of 63; the divergence then carries
// step_div_scale() = 1/(2 h^2) once per cell.
template <typename T>
PF5D_DEV void step_face_flux(const T* Xa, const T* Xb, const T* Ma, const T* Mb, const T* mua, const T* mub,
                             T inv_dx, T* out) {
  if constexpr (sizeof(T) == 8) {
    T g[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) g[k] = (mub[k] - mua[k]) * inv_dx;
    face_flux<T>(Xa, Xb, Ma, Mb, g, out);
  } else {
    T g[4], Xm[4], sM[5], wk[4], dk[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      g[k] = mub[k] - mua[k];
      Xm[k] = T(0.5) * (Xa[k] + Xb[k]);
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) sM[m] = Ma[m] + Mb[m];
    const T w_fe = (T(1) - (((Xm[0] + Xm[1]) + Xm[2]) + Xm[3])) * sM[0];
    T d = Xm[0] * g[0];
#pragma unroll
    for (int k = 1; k < 4; ++k) d = mad<T>(Xm[k], g[k], d);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      wk[k] = Xm[k] * sM[1 + k];
      dk[k] = g[k] - d;
    }
    T S = wk[0] * dk[0];
#pragma unroll
    for (int k = 1; k < 4; ++k) S = mad<T>(wk[k], dk[k], S);
    const T c = mad<T>(w_fe, d, -S);
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = mad<T>(Xm[k], c, wk[k] * dk[k]);
  }
}
// factor that turns a sum of step_face_flux values into dX/dt
template <typename T>
PF5D_DEV T step_div_scale(T inv_dx) {
  return (sizeof(T) == 8) ? inv_dx : T(0.5) * inv_dx * inv_dx;
}

// ---------------------------------------------------------------------------
// Canonical shape: (n0, n1, n2) row-major, the reference's axis order.
// 1D (N) -> (N,1,1) with dim=1; 2D (NX,NY) -> (NX,NY,1) dim=2; 3D as is.
// ---------------------------------------------------------------------------
struct Shape {
  int dim;
  int n[3];
  long ncell;
};

PF5D_DEV long nbr_index(const Shape& sh, long c, int axis, int step) {
  // index of the periodic neighbour of cell c along axis (0..dim-1)
  long stride = 1;
  for (int a = sh.dim - 1; a > axis; --a) stride *= sh.n[a];
  const long i = (c / stride) % sh.n[axis];
  long j = i + step;
  if (j < 0) j += sh.n[axis];
  if (j >= sh.n[axis]) j -= sh.n[axis];
  return c + (j - i) * stride;
}

// interpolator_nb (interpolation.py:154-211): out (ncell, 9), row stride ostride
template <typename T>
__global__ void k_interp(const T* __restrict__ X, long ncell, TableView<T> tv, T* __restrict__ out,
                         long ostride) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  T x[4], P[9];
  V4<T> v = ld4<T>(X + 4 * c);
#pragma unroll
  for (int a = 0; a < 4; ++a) x[a] = v.v[a];
  interp9<T>(tv, x, P);
#pragma unroll
  for (int p = 0; p < 9; ++p) out[c * ostride + p] = P[p];
}

// calculate_laplacian (operators.py:33-81): neighbours summed x+, x-, y+, y-, z+, z-,
// then minus 2*dim*centre, times 1/h^2
template <typename T>
__global__ void k_laplacian(const T* __restrict__ X, Shape sh, T inv_dx2, T* __restrict__ out) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= sh.ncell) return;
  V4<T> xc = ld4<T>(X + 4 * c);
  T acc[4];
  bool first = true;
  for (int a = 0; a < sh.dim; ++a) {
    for (int st = 1; st >= -1; st -= 2) {
      V4<T> xn = ld4<T>(X + 4 * nbr_index(sh, c, a, st));
#pragma unroll
      for (int s = 0; s < 4; ++s) acc[s] = first ? xn.v[s] : acc[s] + xn.v[s];
      first = false;
    }
  }
  const T cdim = T(2 * sh.dim);
  V4<T> o;
#pragma unroll
  for (int s = 0; s < 4; ++s) o.v[s] = (acc[s] - cdim * xc.v[s]) * inv_dx2;
  st4<T>(out + 4 * c, o);
}

// composition_laplacian *= kappa ; mu -= composition_laplacian (simulate.py:254-255)
template <typename T>
__global__ void k_kappa_sub(T* __restrict__ L, T* __restrict__ mu, long mstride, long ncell, V4<T> kappa) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  V4<T> l = ld4<T>(L + 4 * c);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    l.v[s] = l.v[s] * kappa.v[s];
    mu[c * mstride + s] = mu[c * mstride + s] - l.v[s];
  }
  st4<T>(L + 4 * c, l);
}

// _dot_misfit_flat (elastic.py:21-29): s = sum_s (X_s - X0_s) l_s
template <typename T>
__global__ void k_misfit(const T* __restrict__ X, long ncell, V4<T> X0, V4<T> l, T* __restrict__ s) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  V4<T> x = ld4<T>(X + 4 * c);
  T acc = (x.v[0] - X0.v[0]) * l.v[0];
#pragma unroll
  for (int k = 1; k < 4; ++k) acc = mad<T>(x.v[k] - X0.v[k], l.v[k], acc);
  s[c] = acc;
}

// chemical_potentials += psi (x) l (elastic.py:262-263)
template <typename T>
__global__ void k_add_psi(T* __restrict__ mu, long mstride, const T* __restrict__ psi, long ncell, V4<T> l) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const T p = psi[c];
#pragma unroll
  for (int s = 0; s < 4; ++s) mu[c * mstride + s] = mu[c * mstride + s] + p * l.v[s];
}

// _compute_divergence_from_mu_{1,2,3}d (fluxes.py:296-395): every cell evaluates
// its 2*dim faces, as the reference does
template <typename T>
__global__ void k_div_from_mu(const T* __restrict__ X, const T* __restrict__ mob, long bstride,
                              const T* __restrict__ mu, long mstride, Shape sh, T inv_dx, T* __restrict__ out) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= sh.ncell) return;
  T Xc[4], Mc[5], muc[4];
  V4<T> xv = ld4<T>(X + 4 * c);
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    Xc[s] = xv.v[s];
    muc[s] = mu[c * mstride + s];
  }
#pragma unroll
  for (int m = 0; m < 5; ++m) Mc[m] = mob[c * bstride + m];
  T acc[4];
  for (int a = 0; a < sh.dim; ++a) {
    T fp[4], fm[4];
    for (int st = 1; st >= -1; st -= 2) {
      const long n = nbr_index(sh, c, a, st);
      T Xn[4], Mn[5], g[4];
      V4<T> xn = ld4<T>(X + 4 * n);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        Xn[s] = xn.v[s];
        const T mun = mu[n * mstride + s];
        g[s] = (st > 0 ? (mun - muc[s]) : (muc[s] - mun)) * inv_dx;
      }
#pragma unroll
      for (int m = 0; m < 5; ++m) Mn[m] = mob[n * bstride + m];
      face_flux<T>(Xc, Xn, Mc, Mn, g, st > 0 ? fp : fm);
    }
#pragma unroll
    for (int s = 0; s < 4; ++s) acc[s] = (a == 0) ? (fp[s] - fm[s]) : ((acc[s] + fp[s]) - fm[s]);
  }
  V4<T> o;
#pragma unroll
  for (int s = 0; s < 4; ++s) o.v[s] = acc[s] * inv_dx;
  st4<T>(out + 4 * c, o);
}

// calculate_gradients_pm (operators.py:88-158): per axis d, forward
// (A(c+e_d) - A(c)) / h and backward (A(c) - A(c-e_d)) / h differences of a
// 4-component field with row stride astride; out (dim, ncell, 4)
template <typename T>
__global__ void k_gradients_pm(const T* __restrict__ A, long astride, Shape sh, T inv_dx, T* __restrict__ gp,
                               T* __restrict__ gm) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= sh.ncell) return;
  for (int a = 0; a < sh.dim; ++a) {
    const long ip = nbr_index(sh, c, a, 1), im = nbr_index(sh, c, a, -1);
    V4<T> p, m;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const T center = A[c * astride + s];
      p.v[s] = (A[ip * astride + s] - center) * inv_dx;
      m.v[s] = (center - A[im * astride + s]) * inv_dx;
    }
    st4<T>(gp + 4 * ((long)a * sh.ncell + c), p);
    st4<T>(gm + 4 * ((long)a * sh.ncell + c), m);
  }
}

// get_flux_function(dim) (fluxes.py:6-125): the flux through both faces of
// every cell along every axis from the gradients above; out (dim, ncell, 4)
template <typename T>
__global__ void k_fluxes_pm(const T* __restrict__ X, const T* __restrict__ mob, long bstride, Shape sh,
                            const T* __restrict__ gp, const T* __restrict__ gm, T* __restrict__ fp,
                            T* __restrict__ fm) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= sh.ncell) return;
  T Xc[4], Mc[5];
  const V4<T> xv = ld4<T>(X + 4 * c);
#pragma unroll
  for (int s = 0; s < 4; ++s) Xc[s] = xv.v[s];
#pragma unroll
  for (int m = 0; m < 5; ++m) Mc[m] = mob[c * bstride + m];
  for (int a = 0; a < sh.dim; ++a) {
    for (int st = 1; st >= -1; st -= 2) {
      const long n = nbr_index(sh, c, a, st);
      const long o = 4 * ((long)a * sh.ncell + c);
      T Xn[4], Mn[5], g[4], f[4];
      const V4<T> xn = ld4<T>(X + 4 * n), gv = ld4<T>((st > 0 ? gp : gm) + o);
#pragma unroll
      for (int s = 0; s < 4; ++s) {
        Xn[s] = xn.v[s];
        g[s] = gv.v[s];
      }
#pragma unroll
      for (int m = 0; m < 5; ++m) Mn[m] = mob[n * bstride + m];
      face_flux<T>(Xc, Xn, Mc, Mn, g, f);
      V4<T> fv;
#pragma unroll
      for (int s = 0; s < 4; ++s) fv.v[s] = f[s];
      st4<T>((st > 0 ? fp : fm) + o, fv);
    }
  }
}

// Field statistics of an (ncell, ncomp <= 8) array with row stride astride,
// in FP64: per component a Neumaier sum, min and max, and the max of the row
// sums.  acc[0..ncomp) sums, [8..16) compensations, [16..24) min, [24..32)
// max, [32] max row sum; acc is prepared by the caller (k_stats_init).
static __global__ void k_stats_init(double* acc) {
  const int t = threadIdx.x;
  if (t < 16) acc[t] = 0.0;
  else if (t < 24) acc[t] = 1e300;
  else if (t < 33) acc[t] = -1e300;
}
static PF5D_DEV void atomic_min_double(double* a, double v) {
  unsigned long long* p = reinterpret_cast<unsigned long long*>(a);
  unsigned long long old = *p;
  while (v < __longlong_as_double((long long)old)) {
    const unsigned long long prev = atomicCAS(p, old, (unsigned long long)__double_as_longlong(v));
    if (prev == old) break;
    old = prev;
  }
}
static PF5D_DEV void atomic_max_double(double* a, double v) {
  unsigned long long* p = reinterpret_cast<unsigned long long*>(a);
  unsigned long long old = *p;
  while (v > __longlong_as_double((long long)old)) {
    const unsigned long long prev = atomicCAS(p, old, (unsigned long long)__double_as_longlong(v));
    if (prev == old) break;
    old = prev;
  }
}
template <typename T>
__global__ void __launch_bounds__(256) k_stats(const T* __restrict__ A, long ncell, long astride, int ncomp,
                                               double* __restrict__ part, double* __restrict__ acc) {
  __shared__ double red[4][33];
  double s[8], cmp[8], mn[8], mx[8], rmax = -1e300;
  for (int k = 0; k < 8; ++k) {
    s[k] = cmp[k] = 0.0;
    mn[k] = 1e300;
    mx[k] = -1e300;
  }
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (long)gridDim.x * blockDim.x) {
    double rs = 0.0;
    for (int k = 0; k < ncomp; ++k) {
      const double v = (double)A[c * astride + k];
      const double t = s[k] + v;
      cmp[k] += (fabs(s[k]) >= fabs(v)) ? ((s[k] - t) + v) : ((v - t) + s[k]);
      s[k] = t;
      mn[k] = fmin(mn[k], v);
      mx[k] = fmax(mx[k], v);
      rs += v;
    }
    rmax = fmax(rmax, rs);
  }
  // block reduction: plain sums of the per-thread Neumaier sums (the compensation
  // terms are carried separately), then one partial per block
  for (int k = 0; k < ncomp; ++k) {
    double a = s[k], b = cmp[k], lo = mn[k], hi = mx[k];
    for (int off = 32; off > 0; off >>= 1) {
      a += __shfl_xor(a, off, 64);
      b += __shfl_xor(b, off, 64);
      lo = fmin(lo, __shfl_xor(lo, off, 64));
      hi = fmax(hi, __shfl_xor(hi, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
      red[threadIdx.x >> 6][k] = a;
      red[threadIdx.x >> 6][8 + k] = b;
      red[threadIdx.x >> 6][16 + k] = lo;
      red[threadIdx.x >> 6][24 + k] = hi;
    }
  }
  for (int off = 32; off > 0; off >>= 1) rmax = fmax(rmax, __shfl_xor(rmax, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][32] = rmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < ncomp; ++k) {
      double a = 0, b = 0, lo = 1e300, hi = -1e300;
      for (int w = 0; w < 4; ++w) {
        a += red[w][k];
        b += red[w][8 + k];
        lo = fmin(lo, red[w][16 + k]);
        hi = fmax(hi, red[w][24 + k]);
      }
      part[blockIdx.x * 16 + k] = a;
      part[blockIdx.x * 16 + 8 + k] = b;
      atomic_min_double(&acc[16 + k], lo);
      atomic_max_double(&acc[24 + k], hi);
    }
    double r = red[0][32];
    for (int w = 1; w < 4; ++w) r = fmax(r, red[w][32]);
    atomic_max_double(&acc[32], r);
  }
}
// second pass: sum of squared deviations from the given per-component means
template <typename T>
__global__ void __launch_bounds__(256) k_sqdev(const T* __restrict__ A, long ncell, long astride, int ncomp,
                                               const double* __restrict__ mean, double* __restrict__ part) {
  __shared__ double red[4][8];
  double q[8];
  for (int k = 0; k < 8; ++k) q[k] = 0.0;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (long)gridDim.x * blockDim.x)
    for (int k = 0; k < ncomp; ++k) {
      const double d = (double)A[c * astride + k] - mean[k];
      q[k] += d * d;
    }
  for (int k = 0; k < ncomp; ++k) {
    double a = q[k];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][k] = a;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < ncomp; ++k) part[blockIdx.x * 8 + k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
}

// per-component FP64 dot products sum_c A[c,k] B[c,k] of two (ncell, ncomp)
// arrays (contiguous rows), one partial per block
template <typename T>
__global__ void __launch_bounds__(256) k_dot(const T* __restrict__ A, const T* __restrict__ B, long ncell, int ncomp,
                                             double* __restrict__ part) {
  __shared__ double red[4][8];
  double q[8];
  for (int k = 0; k < 8; ++k) q[k] = 0.0;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (long)gridDim.x * blockDim.x)
    for (int k = 0; k < ncomp; ++k) q[k] += (double)A[c * ncomp + k] * (double)B[c * ncomp + k];
  for (int k = 0; k < ncomp; ++k) {
    double a = q[k];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6][k] = a;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int k = 0; k < ncomp; ++k) part[blockIdx.x * 8 + k] = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
}

template <typename T>
__global__ void k_fill(T* p, long n, T v) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) p[c] = v;
}

// abs_max (operators.py:161-175): max |a| ignoring NaN, into *res (zeroed before)
template <typename T>
__global__ void k_absmax(const T* __restrict__ a, long n, T* res) {
  __shared__ T red[16];
  T m = T(0);
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    m = max_ignore_nan<T>(m, fabs(a[i]));
  m = wave_max<T>(m);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    T r = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) r = (red[i] > r) ? red[i] : r;
    atomic_max_nonneg(res, r);
  }
}

// composition_change *= dt ; current_composition += composition_change (simulate.py:316-317)
template <typename T>
__global__ void k_update_plain(T* __restrict__ X, T* __restrict__ dX, long n, double dt) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T d = dX[i] * T(dt);
  dX[i] = d;
  X[i] = X[i] + d;
}

// clamp_simplex (operators.py:262-317)
template <typename T>
__global__ void k_clamp(T* __restrict__ X, long ncell) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  V4<T> x = ld4<T>(X + 4 * c);
  const T tot = ((x.v[0] + x.v[1]) + x.v[2]) + x.v[3];
  if (tot > T(1)) {
    const T inv = T(1) / tot;
#pragma unroll
    for (int s = 0; s < 4; ++s) x.v[s] = x.v[s] * inv;
    st4<T>(X + 4 * c, x);
  }
}

// out_of_bounds statistics (system.py:12-53): per block
// [nonfinite count, max sum, min sum, min comp, max comp]
template <typename T>
__global__ void k_oob_partial(const T* __restrict__ X, long ncell, double* part) {
  __shared__ double red[16][5];
  double nf = 0, smax = -1e300, smin = 1e300, cmin = 1e300, cmax = -1e300;
  for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += (long)gridDim.x * blockDim.x) {
    V4<T> x = ld4<T>(X + 4 * c);
    double s = 0;
    bool fin = true;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double v = (double)x.v[k];
      fin = fin && isfinite(v);
      s += v;
      cmin = fmin(cmin, v);
      cmax = fmax(cmax, v);
    }
    if (!fin) nf += 1;
    smax = fmax(smax, s);
    smin = fmin(smin, s);
  }
  double v[5] = {nf, smax, smin, cmin, cmax};
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v[0] += __shfl_xor(v[0], off, 64);
    v[1] = fmax(v[1], __shfl_xor(v[1], off, 64));
    v[2] = fmin(v[2], __shfl_xor(v[2], off, 64));
    v[3] = fmin(v[3], __shfl_xor(v[3], off, 64));
    v[4] = fmax(v[4], __shfl_xor(v[4], off, 64));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    for (int k = 0; k < 5; ++k) red[w][k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) {
      v[0] += red[i][0];
      v[1] = fmax(v[1], red[i][1]);
      v[2] = fmin(v[2], red[i][2]);
      v[3] = fmin(v[3], red[i][3]);
      v[4] = fmax(v[4], red[i][4]);
    }
    for (int k = 0; k < 5; ++k) part[blockIdx.x * 5 + k] = v[k];
  }
}

}  // namespace pf5d
