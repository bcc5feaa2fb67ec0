// Wave-resident passes of the 3D elastic transform pair for fields whose y and
// z lengths (plane passes) or x length (x pass) are 128.
//
// A 128-point line is transformed by 8 threads of one wavefront, each holding
// 16 points in registers: pass 1 (radix 16, no twiddles) on the points
// t + 8 r (r = 0..15) of thread t, an exchange through a wave-private LDS
// buffer, pass 2 (radix 8, twiddles exp(-2 pi i r k / 128)) on two butterflies
// k = t and t + 8.  Thread t then holds the outputs t + 8 q (q = 0..15),
// which is also the input layout of pass 1: a second transform of the same
// line (the x pass's forward . Bhat . forward) needs no exchange in between.
// A wavefront transforms 8 lines at once and needs no block barrier: only the
// plane passes' transpose between their row and column stages does.
//
// The butterflies (dft16, dft8), the complex multiplies and the twiddle table
// are those of the Stockham passes in pf5d_fft.h (plan 128 = 16 x 8), in the
// same order, so these kernels give the same bits as the line and plane passes
// they replace.
#pragma once
#include "pf5d_fft.h"
#include "pf5d_fftplane.h"

namespace pf5d {

constexpr int FR_N = 128;   // line length
constexpr int FR_NH = 65;   // half spectrum of a 128-point real line
constexpr int FR_LS = 136;  // LDS slots of one line in an exchange buffer
// slot of point e of a line: one spare slot per 16 points, so that the pass-1
// writes (stride 16 points across threads) do not all hit the same banks
PF5D_DEV constexpr int fr_pad(int e) { return e + (e >> 4); }

// orders the wavefront's LDS writes before its reads of other lanes' values
// (a wavefront's LDS instructions complete in order; this keeps the compiler
// from moving them across)
PF5D_DEV void fr_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// the second pass's twiddles of butterflies t and t + 8: plan entries k*7 + r - 1
template <typename T>
struct FrTw {
  cx<T> w[2][7];
};
template <typename T>
PF5D_DEV void fr_load_tw(const cx<T>* __restrict__ tw2, int t, FrTw<T>& W) {
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int r = 1; r < 8; ++r) W.w[h][r - 1] = gld<cx<T>>(tw2 + (t + 8 * h) * 7 + r - 1);
}

// forward FFT of one 128-point line: in, v[r] = x[t + 8 r]; out, v[q] = X[t + 8 q].
// S: this line's FR_LS exchange slots (private to the wavefront)
template <typename T>
PF5D_DEV void fr_fft128(cx<T> v[16], int t, const FrTw<T>& W, cx<T>* S) {
  dft16(v);  // pass 1: butterfly t, outputs y[16 t + r]
#pragma unroll
  for (int r = 0; r < 16; ++r) S[fr_pad(16 * t + r)] = v[r];
  fr_wave_sync();
  cx<T> a[8], b[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    a[r] = S[fr_pad(t + 16 * r)];
    b[r] = S[fr_pad(t + 8 + 16 * r)];
  }
  fr_wave_sync();  // the slots may be written again
#pragma unroll
  for (int r = 1; r < 8; ++r) {
    a[r] = cmul(a[r], W.w[0][r - 1]);
    b[r] = cmul(b[r], W.w[1][r - 1]);
  }
  dft8(a);  // butterfly t: X[t + 16 r]
  dft8(b);  // butterfly t + 8: X[t + 8 + 16 r]
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    v[2 * r] = a[r];
    v[2 * r + 1] = b[r];
  }
}

// LDS of the forward plane pass: the kept spectrum columns (point y at slot
// fr_pad(y)), then NW exchange buffers of 8 lines
template <typename T, int NW, int P>
struct FrPlaneFwd {
  static constexpr int KC = (FR_NH + P - 1) / P;
  static constexpr size_t BYTES = ((size_t)KC * FR_LS + (size_t)NW * 8 * FR_LS) * sizeof(cx<T>);
};
// LDS of the inverse plane pass: the kept rows (point kz at slot fr_pad(kz)),
// then NW exchange buffers
constexpr int FR_LSR = 72;  // slots of one kept row (65 points padded)
template <typename T, int NW, int P>
struct FrPlaneInv {
  static constexpr int RP = (FR_N / 2 + P - 1) / P;  // row pairs of the largest part
  static constexpr size_t BYTES = ((size_t)2 * RP * FR_LSR + (size_t)NW * 8 * FR_LS) * sizeof(cx<T>);
};

// Forward plane pass: rows R2C (every row pair, shared by the parts), then the
// y transform of part p's spectrum columns [k0, k1), written to the spectrum.
template <typename T, int NW, int P>
__global__ void __launch_bounds__(NW * 64) k_frp_fwd(const FftPlaneItem* __restrict__ items, int nitems,
                                                     const LinePlanDev* __restrict__ plans,
                                                     const cx<T>* __restrict__ tw, const SimState* __restrict__ st) {
  using S = FrPlaneFwd<T, NW, P>;
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* PL = reinterpret_cast<cx<T>*>(smem);
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int w = tid >> 6, lane = tid & 63, gl = lane >> 3, t = lane & 7;
  cx<T>* X = PL + S::KC * FR_LS + (w * 8 + gl) * FR_LS;  // this line's exchange slots
  FrTw<T> W;
  int staged = -1;
  int ib, ie;
  block_range(nitems, ib, ie);
  for (int it = ib; it < ie; ++it) {
    const FftPlaneItem I = items[it];
    if (st && !(st[I.sim].status & ST_ACTIVE)) continue;  // uniform per block
    if (I.plan != staged) {
      fr_load_tw<T>(tw + plans[I.plan].tw_off[1], t, W);
      staged = I.plan;
    }
    const int k0 = I.part * FR_NH / P, k1 = (I.part + 1) * FR_NH / P, nk = k1 - k0;
    const T* s = reinterpret_cast<const T*>(I.in) + (long)I.x * FR_N * FR_N;
    // rows: complex line g carries rows 2g (real part) and 2g + 1 (imaginary part)
#pragma unroll 1
    for (int g0 = 0; g0 < FR_N / 2; g0 += NW * 8) {
      const int g = g0 + w * 8 + gl;
      const bool on = g < FR_N / 2;
      cx<T> v[16];
      const T* ra = s + (long)(2 * g) * FR_N + t;
#pragma unroll
      for (int r = 0; r < 16; ++r) v[r] = on ? cx<T>{gld<T>(ra + 8 * r), gld<T>(ra + FR_N + 8 * r)} : cx<T>{T(0), T(0)};
      fr_fft128<T>(v, t, W, X);
      // Z = FFT(a + i b): A[k] = (Z[k] + conj(Z[-k])) / 2, B[k] = -i (Z[k] - conj(Z[-k])) / 2.
      // Z[-k] of k = t + 8 q sits in thread (8 - t) mod 8 at q' = 15 - q (t > 0) or 16 - q (t = 0)
      const int partner = (lane & ~7) | ((8 - t) & 7);
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int m = t + 8 * q;
        cx<T> zr;
        const cx<T> o = v[15 - q];
        zr.x = __shfl(o.x, partner, 64);
        zr.y = __shfl(o.y, partner, 64);
        if (t == 0) zr = v[(16 - q) & 15];
        if (on && m >= k0 && m < k1) {
          const cx<T> z = v[q];
          cx<T>* col = PL + (m - k0) * FR_LS;
          col[fr_pad(2 * g)] = {T(0.5) * (z.x + zr.x), T(0.5) * (z.y - zr.y)};
          col[fr_pad(2 * g + 1)] = {T(0.5) * (z.y + zr.y), T(0.5) * (zr.x - z.x)};
        }
      }
    }
    __syncthreads();
    // y transform of the kept columns; point ky of column k0 + j to C[ky * nhp + k0 + j]
    cx<T>* C = reinterpret_cast<cx<T>*>(I.out) + (long)I.x * FR_N * I.nhp;
    const int nhp = I.nhp;
#pragma unroll 1
    for (int j0 = 0; j0 < nk; j0 += NW * 8) {
      const int j = j0 + w * 8 + gl;
      const bool on = j < nk;
      cx<T> v[16];
      const cx<T>* col = PL + (on ? j : 0) * FR_LS;
#pragma unroll
      for (int r = 0; r < 16; ++r) v[r] = col[fr_pad(t + 8 * r)];
      fr_fft128<T>(v, t, W, X);
      if (on) {
#pragma unroll
        for (int q = 0; q < 16; ++q) gst<cx<T>>(C + (long)(t + 8 * q) * nhp + k0 + j, v[q]);
      }
    }
    __syncthreads();  // PL is reused by the next item
  }
}

// Inverse plane pass: the y transform (inverse: conjugate in, forward,
// conjugate out) of every spectrum column, keeping rows [y0, y1) of part p,
// then those rows' C2R, written to psi.
template <typename T, int NW, int P>
__global__ void __launch_bounds__(NW * 64) k_frp_inv(const FftPlaneItem* __restrict__ items, int nitems,
                                                     const LinePlanDev* __restrict__ plans,
                                                     const cx<T>* __restrict__ tw, const SimState* __restrict__ st) {
  using S = FrPlaneInv<T, NW, P>;
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* PL = reinterpret_cast<cx<T>*>(smem);  // kept row i = y0 + i, point kz at fr_pad(kz)
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int w = tid >> 6, lane = tid & 63, gl = lane >> 3, t = lane & 7;
  cx<T>* X = PL + 2 * S::RP * FR_LSR + (w * 8 + gl) * FR_LS;
  FrTw<T> W;
  int staged = -1;
  int ib, ie;
  block_range(nitems, ib, ie);
  for (int it = ib; it < ie; ++it) {
    const FftPlaneItem I = items[it];
    if (st && !(st[I.sim].status & ST_ACTIVE)) continue;  // uniform per block
    if (I.plan != staged) {
      fr_load_tw<T>(tw + plans[I.plan].tw_off[1], t, W);
      staged = I.plan;
    }
    const int p0 = I.part * (FR_N / 2) / P, p1 = (I.part + 1) * (FR_N / 2) / P, np = p1 - p0;
    const int y0 = 2 * p0, y1 = 2 * p1;
    const cx<T>* C = reinterpret_cast<const cx<T>*>(I.in) + (long)I.x * FR_N * I.nhp;
    const int nhp = I.nhp;
#pragma unroll 1
    for (int c0 = 0; c0 < FR_NH; c0 += NW * 8) {
      const int c = c0 + w * 8 + gl;
      const bool on = c < FR_NH;
      cx<T> v[16];
      const cx<T>* src = C + (on ? c : 0) + (long)t * nhp;
#pragma unroll
      for (int r = 0; r < 16; ++r) v[r] = on ? conjg(gld<cx<T>>(src + (long)(8 * r) * nhp)) : cx<T>{T(0), T(0)};
      fr_fft128<T>(v, t, W, X);
      if (on) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int m = t + 8 * q;
          if (m >= y0 && m < y1) PL[(m - y0) * FR_LSR + fr_pad(c)] = conjg(v[q]);
        }
      }
    }
    __syncthreads();
    // rows C2R: complex line g carries conj(A + i B) of rows y0 + 2g, y0 + 2g + 1,
    // from the Hermitian extension of their half spectra
    T* psi = reinterpret_cast<T*>(I.out) + ((long)I.x * FR_N + y0) * FR_N;
#pragma unroll 1
    for (int g0 = 0; g0 < np; g0 += NW * 8) {
      const int g = g0 + w * 8 + gl;
      const bool on = g < np;
      const cx<T>* La = PL + (on ? 2 * g : 0) * FR_LSR;
      const cx<T>* Lb = La + FR_LSR;
      cx<T> v[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int m = t + 8 * r;
        const cx<T> ca = (m < FR_NH) ? conjg(La[fr_pad(m)]) : La[fr_pad(FR_N - m)];
        const cx<T> cb = (m < FR_NH) ? conjg(Lb[fr_pad(m)]) : Lb[fr_pad(FR_N - m)];
        v[r] = {ca.x + cb.y, ca.y - cb.x};
      }
      fr_fft128<T>(v, t, W, X);
      if (on) {
        T* ra = psi + (long)(2 * g) * FR_N + t;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          gst<T>(ra + 8 * q, v[q].x);
          gst<T>(ra + FR_N + 8 * q, -v[q].y);
        }
      }
    }
    __syncthreads();  // PL is reused by the next item
  }
}

// x pass: lines (ky, kz) along x, point x at C[(x * ny + ky) * nhp + kz]:
// forward, W = conj(Z) Bhat, forward, conj to C.  An item is NW * 8 lines,
// line l = L0 + l of a simulation's ny * nh lines with kz fastest.
struct FrXItem {
  int sim, plan, L0, nl, ny, nh, nhp;
  void* C;
  const void* bhat;
};

// LDS of the x pass: NW exchange buffers
template <typename T, int NW>
struct FrX {
  static constexpr size_t BYTES = (size_t)NW * 8 * FR_LS * sizeof(cx<T>);
};

template <typename T, int NW>
__global__ void __launch_bounds__(NW * 64) k_frx(const FrXItem* __restrict__ items, int nitems,
                                                 const LinePlanDev* __restrict__ plans, const cx<T>* __restrict__ tw,
                                                 const SimState* __restrict__ st) {
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* XB = reinterpret_cast<cx<T>*>(smem);
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int w = tid >> 6, lane = tid & 63, gl = lane >> 3, t = lane & 7;
  cx<T>* X = XB + (w * 8 + gl) * FR_LS;
  FrTw<T> W;
  int staged = -1;
  int ib, ie;
  block_range(nitems, ib, ie);
  for (int it = ib; it < ie; ++it) {
    const FrXItem I = items[it];
    if (st && !(st[I.sim].status & ST_ACTIVE)) continue;  // uniform per block
    if (I.plan != staged) {
      fr_load_tw<T>(tw + plans[I.plan].tw_off[1], t, W);
      staged = I.plan;
    }
    const int l = w * 8 + gl;
    const bool on = l < I.nl;
    const int L = I.L0 + (on ? l : 0);
    const int ky = L / I.nh, kz = L - ky * I.nh;
    const long ps = (long)I.ny * I.nhp;  // point stride along x
    cx<T>* C = reinterpret_cast<cx<T>*>(I.C) + (long)ky * I.nhp + kz + t * ps;
    const T* B = reinterpret_cast<const T*>(I.bhat) + (long)ky * I.nhp + kz + t * ps;
    cx<T> v[16];
    T b[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      v[r] = gld<cx<T>>(C + 8 * r * ps);
      b[r] = gld<T>(B + 8 * r * ps);
    }
    // the two transforms share one copy of the code (a loop, not unrolled): the
    // kernel's instructions are fetched once per wave and run twice
#pragma unroll 1
    for (int pass = 0; pass < 2; ++pass) {
      fr_fft128<T>(v, t, W, X);
      if (pass == 0) {
#pragma unroll
        for (int q = 0; q < 16; ++q) v[q] = {v[q].x * b[q], -v[q].y * b[q]};
      }
    }
    if (on) {
#pragma unroll
      for (int q = 0; q < 16; ++q) gst<cx<T>>(C + 8 * q * ps, conjg(v[q]));
    }
  }
}

}  // namespace pf5d
