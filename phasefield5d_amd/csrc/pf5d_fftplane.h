// Plane passes of the 3D elastic transform pair (y and z lengths both N, a
// length with a compile-time plan).
//
// The line passes move the spectrum through global memory five times per step
// (rows R2C, y forward, x with Bhat, y inverse, rows C2R).  Here the z and y
// transforms of one x-plane run in one work item, with the plane held in LDS
// between them, so a step makes three passes: plane forward, x with Bhat (the
// line pass), plane inverse.
//
// An x-plane has N rows of N real points and NH = N/2 + 1 spectrum columns.
// The work of a plane is split P ways so that there are enough blocks to fill
// the chip:
// - forward, part p: every row pair is transformed (rows R2C; the rows are
//   shared by all parts, so each part redoes them), the spectrum columns of part
//   p are kept in LDS and transformed along y, and their last pass writes the
//   spectrum;
// - inverse, part p: every column is transformed along y (again shared), the
//   rows of part p are kept in LDS and transformed along z (C2R), and written.
// Per line the arithmetic is that of the line passes -- same plans, same
// butterflies, same order -- so both routes give the same bits.
#pragma once
#include "pf5d_fft.h"

namespace pf5d {

struct FftPlaneItem {
  int sim, plan, x, part;
  const void* in;  // forward: s (real, n0 x N x N); inverse: spectrum (rows of nhp)
  void* out;       // forward: spectrum; inverse: psi (real)
  int nhp;         // spectrum row length
};

// LDS of one block on gfx950; the batch sizes below are the largest that fit
constexpr long FFTP_LDS_CAP = 160 * 1024;

template <typename T, int N, int P>
struct PlaneFwd {
  static constexpr int NH = N / 2 + 1, LS = line_stride(N);
  static constexpr int KC = (NH + P - 1) / P;  // spectrum columns of the largest part
  // LDS slots with batches of g row pairs: twiddles, the kept columns, the pass
  // buffer (two batch buffers, or the kept columns' second buffer)
  static constexpr int slots(int g) { return N + KC * LS + ((KC > 2 * g) ? KC : 2 * g) * LS; }
  // row pairs per batch: the largest power of two dividing N/2 that fits
  static constexpr int pick() {
    int g = 1;
    while ((N / 2) % (2 * g) == 0 && (long)slots(2 * g) * (long)sizeof(cx<T>) <= FFTP_LDS_CAP) g *= 2;
    return g;
  }
  static constexpr int GB = pick();
  static constexpr int BL = (KC > 2 * GB) ? KC : 2 * GB;  // lines of the pass buffer
  static constexpr int SLOTS = slots(GB);
};

template <typename T, int N, int P>
struct PlaneInv {
  static constexpr int NH = N / 2 + 1, LS = line_stride(N), LSR = line_stride(NH);
  static constexpr int RP = (N / 2 + P - 1) / P;  // row pairs of the largest part
  static constexpr int NPASS = fft_npass(N);
  // LDS slots with column batches of c lines: twiddles, the kept rows (2 RP rows
  // of NH), the pass buffers (the y stage's one or two of c lines, the rows
  // stage's two of RP lines)
  static constexpr int blines(int c) {
    const int y = (NPASS > 2) ? 2 * c : c;
    return (y > 2 * RP) ? y : 2 * RP;
  }
  static constexpr int slots(int c) { return N + 2 * RP * LSR + blines(c) * LS; }
  // columns per batch: as few batches as fit, evenly sized
  static constexpr int pick() {
    for (int nb = 1; nb < NH; ++nb) {
      const int c = (NH + nb - 1) / nb;
      if ((long)slots(c) * (long)sizeof(cx<T>) <= FFTP_LDS_CAP) return c;
    }
    return 1;
  }
  static constexpr int GC = pick();
  static constexpr int BL = blines(GC);  // lines of the pass buffer region
  static constexpr int SLOTS = slots(GC);
};

template <typename T, int NT>
PF5D_DEV void stage_twiddles(const LinePlanDev& pl, const cx<T>* __restrict__ tw, cx<T>* twl, int tid) {
  for (int i = tid; i < pl.n - 1; i += NT) twl[i] = gld<cx<T>>(tw + pl.tw_off[0] + i);
}

template <typename T, int NT, int N, int P>
__global__ void __launch_bounds__(NT) k_fftp_fwd(const FftPlaneItem* __restrict__ items, int nitems,
                                                 const LinePlanDev* __restrict__ plans,
                                                 const cx<T>* __restrict__ tw, const SimState* __restrict__ st) {
  using S = PlaneFwd<T, N, P>;
  constexpr int LS = S::LS, NH = S::NH, GB = S::GB;
  constexpr int PF = (2 * GB * N + NT - 1) / NT;  // row points per thread per batch
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* twl = reinterpret_cast<cx<T>*>(smem);
  cx<T>* PL = twl + N;          // kept columns: line j = column k0 + j, point y
  cx<T>* B0 = PL + S::KC * LS;  // pass buffer
  int staged = -1;
  int ib, ie;
  block_range(nitems, ib, ie);
  for (int it = ib; it < ie; ++it) {
    const FftPlaneItem I = items[it];
    if (st && !(st[I.sim].status & ST_ACTIVE)) continue;  // uniform per block
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    if (I.plan != staged) {
      stage_twiddles<T, NT>(plans[I.plan], tw, twl, tid);
      staged = I.plan;
      __syncthreads();
    }
    const int k0 = I.part * NH / P, k1 = (I.part + 1) * NH / P, nk = k1 - k0;
    const T* s = reinterpret_cast<const T*>(I.in) + (long)I.x * N * N;
    // rows stage: batches of GB row pairs, the next batch's points loaded
    // into registers while the current one is transformed
    T pre[PF];
    auto load_batch = [&](int r0) {
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int idx = tid + u * NT;
        if (idx < 2 * GB * N) pre[u] = gld<T>(s + (long)r0 * N + idx);
      }
    };
    load_batch(0);
    cx<T>* bA = B0;
    cx<T>* bB = B0 + GB * LS;
    T* bAr = reinterpret_cast<T*>(bA);
#pragma unroll 1
    for (int r0 = 0; r0 < N; r0 += 2 * GB) {
      // complex line g of the batch: rows r0 + 2g (real part), r0 + 2g + 1 (imaginary part)
#pragma unroll
      for (int u = 0; u < PF; ++u) {
        const int idx = tid + u * NT;
        if (idx < 2 * GB * N) {
          const int row = idx / N, m = idx - row * N;
          bAr[2 * ((row >> 1) * LS + lpad(m)) + (row & 1)] = pre[u];
        }
      }
      __syncthreads();
      if (r0 + 2 * GB < N) load_batch(r0 + 2 * GB);
      auto fetch = [&](int g, int m) -> cx<T> { return bA[g * LS + lpad(m)]; };
      auto natural = [&](int g, int m, const cx<T>& v, cx<T>* other) { other[g * LS + lpad(m)] = v; };
      CtRun<T, NT, N, 0, 1>::run(GB, frcp(GB), tid, twl, bB, bA, false, fetch, false, natural);
      __syncthreads();
      const cx<T>* res = (fft_npass(N) % 2 == 1) ? bB : bA;
      // Z = FFT(a + i b): A[k] = (Z[k] + conj(Z[-k])) / 2, B[k] = -i (Z[k] - conj(Z[-k])) / 2
#pragma unroll 1
      for (int idx = tid; idx < 2 * GB * nk; idx += NT) {
        const int row = idx / nk, j = idx - row * nk, m = k0 + j;
        const cx<T>* Z = res + (row >> 1) * LS;
        const cx<T> z = Z[lpad(m)], zr = Z[lpad(m == 0 ? 0 : N - m)];
        const cx<T> o = (row & 1) ? cx<T>{T(0.5) * (z.y + zr.y), T(0.5) * (zr.x - z.x)}
                                  : cx<T>{T(0.5) * (z.x + zr.x), T(0.5) * (z.y - zr.y)};
        PL[j * LS + lpad(r0 + row)] = o;
      }
      __syncthreads();
    }
    // y stage on the kept columns: first pass from PL, last pass to the
    // spectrum with the columns fastest across threads
    cx<T>* C = reinterpret_cast<cx<T>*>(I.out) + (long)I.x * N * I.nhp;
    const int nhp = I.nhp;
    auto fetch_pl = [&](int g, int m) -> cx<T> { return PL[g * LS + lpad(m)]; };
    auto sink_c = [&](int g, int m, const cx<T>& v, cx<T>*) { gst<cx<T>>(C + (long)m * nhp + k0 + g, v); };
    CtRun<T, NT, N, 0, 1>::run(nk, frcp(nk), tid, twl, B0, PL, false, fetch_pl, true, sink_c);
    __syncthreads();
  }
}

template <typename T, int NT, int N, int P>
__global__ void __launch_bounds__(NT) k_fftp_inv(const FftPlaneItem* __restrict__ items, int nitems,
                                                 const LinePlanDev* __restrict__ plans,
                                                 const cx<T>* __restrict__ tw, const SimState* __restrict__ st) {
  using S = PlaneInv<T, N, P>;
  constexpr int LS = S::LS, LSR = S::LSR, NH = S::NH, GC = S::GC;
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* twl = reinterpret_cast<cx<T>*>(smem);
  cx<T>* PL = twl + N;               // kept rows: row i = y0 + i, point kz
  cx<T>* B0 = PL + 2 * S::RP * LSR;  // pass buffers
  // the y stage's second buffer (plans of three or more passes), the rows stage's second buffer
  cx<T>* B1y = B0 + GC * LS;
  cx<T>* B1r = B0 + S::RP * LS;
  int staged = -1;
  int ib, ie;
  block_range(nitems, ib, ie);
  for (int it = ib; it < ie; ++it) {
    const FftPlaneItem I = items[it];
    if (st && !(st[I.sim].status & ST_ACTIVE)) continue;  // uniform per block
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    if (I.plan != staged) {
      stage_twiddles<T, NT>(plans[I.plan], tw, twl, tid);
      staged = I.plan;
      __syncthreads();
    }
    const int p0 = I.part * (N / 2) / P, p1 = (I.part + 1) * (N / 2) / P, np = p1 - p0;
    const int y0 = 2 * p0, y1 = 2 * p1;
    const cx<T>* C = reinterpret_cast<const cx<T>*>(I.in) + (long)I.x * N * I.nhp;
    const int nhp = I.nhp;
    // y stage (inverse: conjugate in, forward transform, conjugate out) on
    // batches of columns, first pass from the spectrum with the columns fastest
#pragma unroll 1
    for (int c0 = 0; c0 < NH; c0 += GC) {
      const int gc = min(GC, NH - c0);
      auto fetch_c = [&](int g, int m) -> cx<T> { return conjg(gld<cx<T>>(C + (long)m * nhp + c0 + g)); };
      auto sink_pl = [&](int g, int m, const cx<T>& v, cx<T>*) {
        if (m >= y0 && m < y1) PL[(m - y0) * LSR + lpad(c0 + g)] = conjg(v);
      };
      CtRun<T, NT, N, 0, 1>::run(gc, frcp(gc), tid, twl, B0, B1y, true, fetch_c, false, sink_pl);
      __syncthreads();
    }
    // rows stage (C2R) on the kept row pairs: complex line g carries
    // conj(A + i B) for rows y0 + 2g, y0 + 2g + 1, from the Hermitian extension
    auto fetch_pl = [&](int g, int m) -> cx<T> {
      auto cfull = [&](int l) -> cx<T> {
        const cx<T>* L = PL + l * LSR;
        return (m < NH) ? conjg(L[lpad(m)]) : L[lpad(N - m)];
      };
      const cx<T> ca = cfull(2 * g), cb = cfull(2 * g + 1);
      return {ca.x + cb.y, ca.y - cb.x};
    };
    auto natural = [&](int g, int m, const cx<T>& v, cx<T>* other) { other[g * LS + lpad(m)] = v; };
    CtRun<T, NT, N, 0, 1>::run(np, frcp(np), tid, twl, B0, B1r, false, fetch_pl, false, natural);
    __syncthreads();
    const cx<T>* res = (fft_npass(N) % 2 == 1) ? B0 : B1r;
    // C2R: Re to row y0 + 2g, -Im to row y0 + 2g + 1, point-contiguous
    T* psi = reinterpret_cast<T*>(I.out) + ((long)I.x * N + y0) * N;
#pragma unroll 1
    for (int idx = tid; idx < np * N; idx += NT) {
      const int g = idx / N, m = idx - g * N;
      const cx<T> v = res[g * LS + lpad(m)];
      gst<T>(psi + (long)(2 * g) * N + m, v.x);
      gst<T>(psi + (long)(2 * g + 1) * N + m, -v.y);
    }
    __syncthreads();
  }
}

}  // namespace pf5d
