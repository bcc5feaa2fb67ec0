// Batched mixed-radix FFT in LDS for the rank-1 elastic term
//     psi = irfftn( Bhat * rfftn(s) )     (phasefield5d/solver/elastic.py:246-263)
//
// Algorithm: Stockham autosort FFT in the formulation of Govindaraju, Lloyd,
// Dotsenko, Smith, Manferdelli, "High performance discrete Fourier transforms
// on graphics processors", SC08: for each pass with radix R and sub-transform
// size Ns, butterfly j reads x[j + r n/R], twiddles by exp(-2 pi i r (j mod Ns)
// / (Ns R)), applies a radix-R DFT and writes y[(j/Ns) Ns R + (j mod Ns) + r Ns].
// Radices 2, 3, 4, 5, 8, 16 are hand-written butterflies; any other factor uses
// a direct radix-R DFT from a per-plan root table.  Twiddles are computed on
// the host in extended precision and rounded to the working type.  Plans split
// the power-of-two part of n into as few passes of radix <= 16 as possible,
// balanced (128 = 16 x 8, 512 = 8 x 8 x 8), then the odd factors.
//
// Every work item transforms G lines of length n.  Point m of line g sits at
// g*ls + m*ps (ps = 1 for rows along the contiguous axis, ls = 1 for pencils
// along a strided axis, so global loads are coalesced either way).
//
// Lengths with a compile-time plan gather their lines into LDS (point-major
// along a strided axis, so the loads coalesce), run every pass with all
// indices constant, and write strided lines straight from the last pass; other
// lengths run every pass through LDS with the plan read at run time.
//
// Inverse transforms use IFFT(x) = conj(FFT(conj(x))).  The 1/N of irfftn is
// folded into Bhat on the host.  Because Bhat is real and even in k, the
// product with the half spectrum can be applied to a full complex line and the
// result's real part equals irfftn's output.
#pragma once
#include "pf5d_common.h"

namespace pf5d {

template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T x, y;
};
template <typename T>
PF5D_DEV cx<T> cadd(cx<T> a, cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
PF5D_DEV cx<T> csub(cx<T> a, cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
PF5D_DEV T fused_mad(T a, T b, T c) {
  if constexpr (sizeof(T) == 4)
    return __builtin_fmaf(a, b, c);
  else
    return __builtin_fma(a, b, c);
}
// two multiplies and two fused multiply-adds
template <typename T>
PF5D_DEV cx<T> cmul(cx<T> a, cx<T> b) {
  return {fused_mad<T>(a.x, b.x, -(a.y * b.y)), fused_mad<T>(a.x, b.y, a.y * b.x)};
}
template <typename T>
PF5D_DEV cx<T> cscale(cx<T> a, T s) { return {a.x * s, a.y * s}; }
template <typename T>
PF5D_DEV cx<T> mul_minus_i(cx<T> a) { return {a.y, -a.x}; }  // -i * a
template <typename T>
PF5D_DEV cx<T> conjg(cx<T> a) { return {a.x, -a.y}; }

constexpr int FFT_MAXPASS = 24;
struct LinePlanDev {
  int n, npass;
  int radix[FFT_MAXPASS];
  int tw_off[FFT_MAXPASS];    // twiddles of pass p: tw[tw_off + k*(R-1) + (r-1)], k < Ns, 1 <= r < R
  int root_off[FFT_MAXPASS];  // roots exp(-2 pi i m / R), m < R (generic radices only)
  int nroot;                  // roots of all generic passes (contiguous from root_off[0])
  int tw_global;              // twiddles and roots stay in global memory (lines too long for
                              // them and two line buffers in LDS)
};

enum FftKind { FK_R2C = 0, FK_FWD = 1, FK_CONV = 2, FK_INV = 3, FK_C2R = 4, FK_LINE = 5 };

struct FftItem {
  int kind, plan, G, sim;
  const void* in;
  void* out;
  const void* bhat;
  long ps_in, ls_in, ps_out, ls_out, ps_b, ls_b;
};

template <typename T>
PF5D_DEV void dft2(cx<T>* v) {
  const cx<T> a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}
template <typename T>
PF5D_DEV void dft4(cx<T>& a0, cx<T>& a1, cx<T>& a2, cx<T>& a3) {
  const cx<T> t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = mul_minus_i(csub(a1, a3));
  a0 = cadd(t0, t2);
  a2 = csub(t0, t2);
  a1 = cadd(t1, t3);
  a3 = csub(t1, t3);
}
template <typename T>
PF5D_DEV void dft8(cx<T>* v) {
  cx<T> e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
  cx<T> o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
  dft4(e0, e1, e2, e3);
  dft4(o0, o1, o2, o3);
  const T c = T(0.70710678118654752440084436210484903928);
  o1 = {c * (o1.x + o1.y), c * (o1.y - o1.x)};    // * (c, -c)
  o2 = mul_minus_i(o2);                           // * (0, -1)
  o3 = {c * (o3.y - o3.x), -c * (o3.x + o3.y)};   // * (-c, -c)
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1);
  v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2);
  v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3);
  v[7] = csub(e3, o3);
}
template <typename T>
PF5D_DEV void dft3(cx<T>* v) {
  const T h = T(0.86602540378443864676372317075293618347);  // sqrt(3)/2
  const cx<T> t = cadd(v[1], v[2]);
  const cx<T> m = {v[0].x - T(0.5) * t.x, v[0].y - T(0.5) * t.y};
  const cx<T> d0 = csub(v[1], v[2]);
  const cx<T> d = {h * d0.y, -h * d0.x};  // -i sqrt(3)/2 (a1 - a2)
  v[0] = cadd(v[0], t);
  v[1] = cadd(m, d);
  v[2] = csub(m, d);
}
template <typename T>
PF5D_DEV void dft5(cx<T>* v) {
  const T c1 = T(0.30901699437494742410229341718281905886);   // cos(2pi/5)
  const T c2 = T(-0.80901699437494742410229341718281905886);  // cos(4pi/5)
  const T s1 = T(0.95105651629515357211643933337938214340);   // sin(2pi/5)
  const T s2 = T(0.58778525229247312916870595463907276860);   // sin(4pi/5)
  const cx<T> t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]), t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
  const cx<T> b1 = {v[0].x + c1 * t1.x + c2 * t2.x, v[0].y + c1 * t1.y + c2 * t2.y};
  const cx<T> b2 = {v[0].x + c2 * t1.x + c1 * t2.x, v[0].y + c2 * t1.y + c1 * t2.y};
  const cx<T> d1 = {s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y};
  const cx<T> d2 = {s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y};
  v[0] = cadd(v[0], cadd(t1, t2));
  const cx<T> md1 = mul_minus_i(d1), md2 = mul_minus_i(d2);
  v[1] = cadd(b1, md1);
  v[4] = csub(b1, md1);
  v[2] = cadd(b2, md2);
  v[3] = csub(b2, md2);
}
// 16 = 4 x 4: X[k1 + 4 k2] = sum_n2 W4^(n2 k2) W16^(n2 k1) sum_n1 x[4 n1 + n2] W4^(n1 k1)
template <typename T>
PF5D_DEV void dft16(cx<T>* v) {
  const T c1 = T(0.92387953251128675612818318939678828682);  // cos(pi/8)
  const T s1 = T(0.38268343236508977172845998403039886676);  // sin(pi/8)
  const T r2 = T(0.70710678118654752440084436210484903928);
  cx<T> y[16];
#pragma unroll
  for (int n2 = 0; n2 < 4; ++n2) {
    cx<T> a0 = v[n2], a1 = v[4 + n2], a2 = v[8 + n2], a3 = v[12 + n2];
    dft4(a0, a1, a2, a3);
    y[n2 * 4 + 0] = a0;
    y[n2 * 4 + 1] = a1;
    y[n2 * 4 + 2] = a2;
    y[n2 * 4 + 3] = a3;
  }
  // twiddles W16^(n2 k1) = exp(-2 pi i n2 k1 / 16)
  y[5] = cmul(y[5], cx<T>{c1, -s1});    // W^1
  y[6] = cmul(y[6], cx<T>{r2, -r2});    // W^2
  y[7] = cmul(y[7], cx<T>{s1, -c1});    // W^3
  y[9] = cmul(y[9], cx<T>{r2, -r2});    // W^2
  y[10] = mul_minus_i(y[10]);           // W^4
  y[11] = cmul(y[11], cx<T>{-r2, -r2});  // W^6
  y[13] = cmul(y[13], cx<T>{s1, -c1});  // W^3
  y[14] = cmul(y[14], cx<T>{-r2, -r2});  // W^6
  y[15] = cmul(y[15], cx<T>{-c1, s1});  // W^9
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    cx<T> b0 = y[k1], b1 = y[4 + k1], b2 = y[8 + k1], b3 = y[12 + k1];
    dft4(b0, b1, b2, b3);
    v[k1] = b0;
    v[k1 + 4] = b1;
    v[k1 + 8] = b2;
    v[k1 + 12] = b3;
  }
}

// Plan: radix of pass p for length n (0 past the last pass).  The power-of-two
// part 2^k takes ceil(k/4) passes with the bits spread evenly (larger radices
// first), then the odd prime factors in increasing order.
PF5D_HD constexpr int fft_radix(int n, int p) {
  int k = 0, m = n;
  while (m % 2 == 0 && m > 1) {
    m /= 2;
    ++k;
  }
  const int P2 = (k + 3) / 4;
  if (p < P2) return 1 << (k / P2 + (p < k % P2 ? 1 : 0));
  int i = P2;
  for (int q = 3; m > 1; q += 2) {
    while (m % q == 0) {
      if (i == p) return q;
      m /= q;
      ++i;
    }
  }
  return 0;
}
PF5D_HD constexpr int fft_npass(int n) {
  int p = 0;
  while (fft_radix(n, p) != 0) ++p;
  return p;
}
PF5D_HD constexpr int fft_min_radix(int n) {
  int r = 1 << 30;
  for (int p = 0; p < fft_npass(n); ++p) r = fft_radix(n, p) < r ? fft_radix(n, p) : r;
  return r;
}

// LDS position of point e of a line: one spare slot after every 8 points, so
// the power-of-two strides of the Stockham reads and writes spread over the
// banks; a line of n points takes line_stride(n) slots
PF5D_DEV constexpr int lpad(int e) { return e + (e >> 3); }
PF5D_HD constexpr int line_stride(int n) { return n + (n >> 3) + 1; }

// radices with hand-written butterflies; every other factor runs the direct DFT
PF5D_HD constexpr bool fft_generic_radix(int R) { return !(R == 2 || R == 3 || R == 4 || R == 5 || R == 8 || R == 16); }
// LDS slots ahead of an item's line buffers: n for the twiddles (n - 1 used),
// then the roots of unity of the generic passes
PF5D_HD constexpr int fft_head_slots(int n) {
  int s = n;
  for (int p = 0; fft_radix(n, p) != 0; ++p)
    if (fft_generic_radix(fft_radix(n, p))) s += fft_radix(n, p);
  return s;
}

// direct radix-R DFT pass, one output point per thread:
//     y_q = x_0 + sum_{p >= 1} (x_p w_p) root^(p q mod R)
// The twiddled inputs x_p w_p are formed once, in place in src (the pass's
// source is dead after it), then every output sums over them; a butterfly's
// outputs sit on adjacent threads, so their LDS reads of x_p coincide.  Same
// operations and order per output as forming x_p w_p inside the sum.
template <typename T, int NT>
PF5D_DEV void generic_pass(cx<T>* __restrict__ src, cx<T>* __restrict__ dst, int n, int LS, int G, int R, int Ns,
                           const cx<T>* __restrict__ tw, const cx<T>* __restrict__ roots) {
  const int nb = n / R, tot = G * n;
  const float rn = frcp(n), rnb = frcp(nb), rns = frcp(Ns), rr = frcp(R);
  for (int u = threadIdx.x; u < tot; u += NT) {
    int g, m, p, j, q, k;
    fdivmod(u, n, rn, g, m);
    fdivmod(m, nb, rnb, p, j);
    if (p == 0) continue;
    fdivmod(j, Ns, rns, q, k);
    cx<T>& x = src[g * LS + lpad(m)];
    x = cmul(x, tw[k * (R - 1) + p - 1]);
  }
  __syncthreads();
  for (int u = threadIdx.x; u < tot; u += NT) {
    int b, qq, g, j, q, k;
    fdivmod(u, R, rr, b, qq);
    fdivmod(b, nb, rnb, g, j);
    fdivmod(j, Ns, rns, q, k);
    const cx<T>* s = src + g * LS;
    cx<T> acc = s[lpad(j)];
    int e = 0;
#pragma unroll 4
    for (int p = 1; p < R; ++p) {
      e += qq;
      if (e >= R) e -= R;
      acc = cadd(acc, cmul(s[lpad(j + p * nb)], roots[e]));
    }
    dst[g * LS + lpad(q * Ns * R + k + qq * Ns)] = acc;
  }
}

// one Stockham pass over G lines (line stride LS): src -> dst.  tw holds this
// pass's twiddles, roots the generic radix's roots of unity (both in LDS).
template <typename T, int NT>
PF5D_DEV void stockham_pass(cx<T>* __restrict__ src, cx<T>* __restrict__ dst, int n, int LS, int G, int R, int Ns,
                            const cx<T>* __restrict__ tw, const cx<T>* __restrict__ roots) {
  if (fft_generic_radix(R)) {
    generic_pass<T, NT>(src, dst, n, LS, G, R, Ns, tw, roots);
    return;
  }
  const int nb = n / R;
  const int tot = G * nb;
  const float rnb = frcp(nb), rns = frcp(Ns);
  for (int b = threadIdx.x; b < tot; b += NT) {
    int g, j, q, k;
    fdivmod(b, nb, rnb, g, j);
    fdivmod(j, Ns, rns, q, k);
    const cx<T>* s = src + g * LS;
    cx<T>* d = dst + g * LS;
    const int ob = q * Ns * R + k;
    const cx<T>* w = tw + k * (R - 1);  // w[r-1] is the twiddle of input r >= 1
    if (R == 16) {
      cx<T> v[16];
      v[0] = s[lpad(j)];
#pragma unroll
      for (int r = 1; r < 16; ++r) v[r] = (Ns == 1) ? s[lpad(j + r * nb)] : cmul(s[lpad(j + r * nb)], w[r - 1]);
      dft16(v);
#pragma unroll
      for (int r = 0; r < 16; ++r) d[lpad(ob + r * Ns)] = v[r];
    } else {
      cx<T> v[8];
      v[0] = s[lpad(j)];
#pragma unroll
      for (int r = 1; r < 8; ++r)
        if (r < R) v[r] = (Ns == 1) ? s[lpad(j + r * nb)] : cmul(s[lpad(j + r * nb)], w[r - 1]);
      if (R == 2)
        dft2(v);
      else if (R == 3)
        dft3(v);
      else if (R == 4)
        dft4(v[0], v[1], v[2], v[3]);
      else if (R == 5)
        dft5(v);
      else
        dft8(v);
#pragma unroll
      for (int r = 0; r < 8; ++r)
        if (r < R) d[lpad(ob + r * Ns)] = v[r];
    }
  }
}

// full forward FFT of G lines held in buf0; returns the buffer holding the result
template <typename T, int NT>
PF5D_DEV cx<T>* fft_lines(cx<T>* buf0, cx<T>* buf1, const LinePlanDev& pl, int LS, int G, const cx<T>* tw,
                          const cx<T>* roots) {
  cx<T>* a = buf0;
  cx<T>* b = buf1;
  int Ns = 1;
  for (int p = 0; p < pl.npass; ++p) {
    const int R = pl.radix[p];
    stockham_pass<T, NT>(a, b, pl.n, LS, G, R, Ns, tw + (pl.tw_off[p] - pl.tw_off[0]),
                         roots + (pl.root_off[p] - pl.root_off[0]));
    __syncthreads();
    cx<T>* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// ---------------------------------------------------------------------------
// Item input and output, shared by both paths
// ---------------------------------------------------------------------------
// Point m of complex line g of an item as the forward transform consumes it:
// R2C packs real lines 2g, 2g+1 as real and imaginary parts; C2R builds
// conj(A + i B) from the Hermitian extensions of half spectra 2g, 2g+1
// (conj(A_full[m]) = conj(A[m]) for m < nh, A[n - m] above); INV conjugates.
template <typename T>
PF5D_DEV cx<T> item_input(const FftItem& I, int G, int n, int nh, int g, int m) {
  if (I.kind == FK_R2C) {
    const T* in = reinterpret_cast<const T*>(I.in);
    const int g0 = 2 * g, g1 = 2 * g + 1;
    return {gld<T>(in + (g0 * I.ls_in + (long)m * I.ps_in)),
            g1 < G ? gld<T>(in + (g1 * I.ls_in + (long)m * I.ps_in)) : T(0)};
  }
  if (I.kind == FK_C2R) {
    const cx<T>* C = reinterpret_cast<const cx<T>*>(I.in);
    auto cfull = [&](int l) -> cx<T> {
      const cx<T>* L = C + l * I.ls_in;
      return (m < nh) ? conjg(gld<cx<T>>(L + (long)m * I.ps_in)) : gld<cx<T>>(L + (long)(n - m) * I.ps_in);
    };
    const cx<T> ca = cfull(2 * g);
    const cx<T> cb = (2 * g + 1 < G) ? cfull(2 * g + 1) : cx<T>{T(0), T(0)};
    return {ca.x + cb.y, ca.y - cb.x};
  }
  if (I.kind == FK_LINE) return {gld<T>(reinterpret_cast<const T*>(I.in) + (g * I.ls_in + (long)m * I.ps_in)), T(0)};
  const cx<T> v = gld<cx<T>>(reinterpret_cast<const cx<T>*>(I.in) + (g * I.ls_in + (long)m * I.ps_in));
  return (I.kind == FK_INV) ? conjg(v) : v;
}

// Bhat factor of point m of line g (CONV and LINE items)
template <typename T>
PF5D_DEV T item_bhat(const FftItem& I, int n, int nh, int g, int m) {
  if (I.kind == FK_LINE) return gld<T>(reinterpret_cast<const T*>(I.bhat) + ((m < nh) ? m : n - m));
  return gld<T>(reinterpret_cast<const T*>(I.bhat) + (g * I.ls_b + (long)m * I.ps_b));
}

// Final point m of complex line g (not R2C): C2R writes Re to real line 2g and
// -Im to line 2g+1 (z = conj(W)); LINE the real part; FWD as is; INV and CONV
// the conjugate.
template <typename T>
PF5D_DEV void item_output(const FftItem& I, int G, int g, int m, const cx<T>& v) {
  if (I.kind == FK_C2R) {
    T* out = reinterpret_cast<T*>(I.out);
    gst<T>(out + (2 * g * I.ls_out + (long)m * I.ps_out), v.x);
    if (2 * g + 1 < G) gst<T>(out + ((2 * g + 1) * I.ls_out + (long)m * I.ps_out), -v.y);
  } else if (I.kind == FK_LINE) {
    gst<T>(reinterpret_cast<T*>(I.out) + (g * I.ls_out + (long)m * I.ps_out), v.x);
  } else {
    const cx<T> o = (I.kind == FK_FWD) ? v : conjg(v);
    gst<cx<T>>(reinterpret_cast<cx<T>*>(I.out) + (g * I.ls_out + (long)m * I.ps_out), o);
  }
}

// ---------------------------------------------------------------------------
// Compile-time path: lengths whose plan is known when the kernel is built.
// Two LDS buffers.  Along a strided axis (consecutive lines adjacent in
// memory) the first pass reads global memory and the last pass writes it
// directly, with the lines varying fastest across threads; rows along the
// contiguous axis are gathered into LDS and stored from it, point-contiguous.
// ---------------------------------------------------------------------------
PF5D_HD constexpr bool ct_length(int n) {
  return n == 64 || n == 96 || n == 128 || n == 192 || n == 256 || n == 512;
}

template <int R, typename T>
PF5D_DEV void dft_r(cx<T>* v) {
  if constexpr (R == 2)
    dft2(v);
  else if constexpr (R == 3)
    dft3(v);
  else if constexpr (R == 4)
    dft4(v[0], v[1], v[2], v[3]);
  else if constexpr (R == 5)
    dft5(v);
  else if constexpr (R == 8)
    dft8(v);
  else
    dft16(v);
}

// One radix-R pass after sub-transforms of size NS over GL lines of N points.
// Butterfly b maps to (line g, index j) line-major, or point-major (pm: the
// lines vary fastest, for global accesses along a strided axis).  Inputs
// src(g, j + r N/R); outputs dst(g, (j/NS) NS R + j%NS + r NS, value).
template <typename T, int NT, int N, int R, int NS, typename Src, typename Dst>
PF5D_DEV void ct_pass(int GL, float rgl, int tid, const cx<T>* __restrict__ tw, bool pm, Src src, Dst dst) {
  constexpr int NB = N / R;
  const int tot = GL * NB;
#pragma unroll 1
  for (int b = tid; b < tot; b += NT) {
    int g, j;
    if (pm) {
      fdivmod(b, GL, rgl, j, g);
    } else {
      g = b / NB;
      j = b - g * NB;
    }
    cx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src(g, j + r * NB);
    const int k = j % NS;
    if constexpr (NS > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[k * (R - 1) + r - 1]);
    }
    dft_r<R>(v);
    const int ob = (j / NS) * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst(g, ob + r * NS, v[r]);
  }
}

// Passes P.. of the plan for N, alternating between buffers b0 and b1: pass 0
// reads fetch and writes b0, pass p > 0 reads the buffer pass p-1 wrote; the
// last pass hands its outputs to sink(g, m, v, other) where other is the
// buffer it does not read.  Pass P's twiddles start at NS - 1 in tw.
template <typename T, int NT, int N, int P, int NS>
struct CtRun {
  template <typename Fetch, typename Sink>
  static PF5D_DEV void run(int GL, float rgl, int tid, const cx<T>* tw, cx<T>* b0, cx<T>* b1, bool pm_src,
                           Fetch fetch, bool pm_dst, Sink sink) {
    constexpr int NP = fft_npass(N), R = fft_radix(N, P), LS = line_stride(N);
    cx<T>* in = (P % 2 == 1) ? b0 : b1;   // written by pass P-1
    cx<T>* out = (P % 2 == 0) ? b0 : b1;  // written by this pass
    auto lds_src = [&](int g, int m) -> cx<T> { return in[g * LS + lpad(m)]; };
    auto lds_dst = [&](int g, int m, const cx<T>& v) { out[g * LS + lpad(m)] = v; };
    const cx<T>* w = tw + (NS - 1);
    if constexpr (P == NP - 1) {
      auto last = [&](int g, int m, const cx<T>& v) { sink(g, m, v, out); };
      if constexpr (P == 0)
        ct_pass<T, NT, N, R, NS>(GL, rgl, tid, w, pm_dst, fetch, last);
      else
        ct_pass<T, NT, N, R, NS>(GL, rgl, tid, w, pm_dst, lds_src, last);
    } else {
      if constexpr (P == 0)
        ct_pass<T, NT, N, R, NS>(GL, rgl, tid, w, pm_src, fetch, lds_dst);
      else
        ct_pass<T, NT, N, R, NS>(GL, rgl, tid, w, false, lds_src, lds_dst);
      __syncthreads();
      CtRun<T, NT, N, P + 1, NS * R>::run(GL, rgl, tid, tw, b0, b1, pm_src, fetch, pm_dst, sink);
    }
  }
};

// dst[slot(idx)] = fetch(idx) for idx = tid, tid + NT, ... < tot, FFT_LB at a
// time: all fetches of a batch are issued before the first LDS write
constexpr int FFT_LB = 4;
// lengths with a compile-time plan (the cases of k_fft's dispatch)
PF5D_HD constexpr bool fft_ct_length(int n) {
  return n == 64 || n == 96 || n == 128 || n == 192 || n == 256 || n == 512;
}
template <typename T, int NT, typename F>
PF5D_DEV void gather_lds(int tot, int tid, cx<T>* dst, F fetch) {
#pragma unroll 1
  for (int base = tid; base < tot; base += NT * FFT_LB) {
    cx<T> v[FFT_LB];
    int at[FFT_LB];
#pragma unroll
    for (int u = 0; u < FFT_LB; ++u) {
      at[u] = -1;
      if (base + u * NT < tot) v[u] = fetch(base + u * NT, at[u]);
    }
#pragma unroll
    for (int u = 0; u < FFT_LB; ++u)
      if (at[u] >= 0) dst[at[u]] = v[u];
  }
}

// LDS: the plan's twiddles (N slots), then two buffers of GL lines (then, for
// CONV items, their GL * N Bhat values)
template <typename T, int NT, int N>
PF5D_DEV void fft_item_ct(const FftItem& I, const cx<T>* twl, cx<T>* b0, int tid) {
  constexpr int LS = line_stride(N), NH = N / 2 + 1;
  const int G = I.G;
  const bool paired = (I.kind == FK_R2C || I.kind == FK_C2R);
  const int GL = paired ? (G + 1) / 2 : G;
  cx<T>* b1 = b0 + GL * LS;
  const float rgl = frcp(GL);
  // strided lines: gathered point-major below; written to global memory from the last pass
  const bool fuse_in = I.ps_in != 1, fuse_out = I.ps_out != 1;
  // rows: gathered point-contiguous into b1 first (the first pass writes b0)
  auto fetch_b1 = [&](int g, int m) -> cx<T> { return b1[g * LS + lpad(m)]; };
  // lines along a strided axis: gathered point-major (lines fastest, so the
  // loads coalesce) into b1 by every thread, then every pass runs from LDS.
  // (Reading them in the first pass instead, with only its butterflies' threads
  // loading, was slower: the x pass at 128^3 took 4 us more in FP32, 9 in FP64.)
  // CONV items gather their Bhat with the lines, into GL * N values after the
  // two line buffers (slot m GL + g), for the first transform's last pass
  const bool conv_bl = fuse_in && I.kind == FK_CONV;
  T* bl = reinterpret_cast<T*>(b1 + GL * LS);
  if (conv_bl) {
    const int tot = GL * N;
#pragma unroll 1
    for (int base = tid; base < tot; base += NT * FFT_LB) {
      cx<T> v[FFT_LB];
      T bb[FFT_LB];
      int at[FFT_LB];
#pragma unroll
      for (int u = 0; u < FFT_LB; ++u) {
        const int idx = base + u * NT;
        at[u] = -1;
        if (idx < tot) {
          int m, g;
          fdivmod(idx, GL, rgl, m, g);
          at[u] = g * LS + lpad(m);
          v[u] = item_input<T>(I, G, N, NH, g, m);
          bb[u] = item_bhat<T>(I, N, NH, g, m);
        }
      }
#pragma unroll
      for (int u = 0; u < FFT_LB; ++u)
        if (at[u] >= 0) {
          b1[at[u]] = v[u];
          bl[base + u * NT] = bb[u];
        }
    }
    __syncthreads();
  } else if (fuse_in) {
    gather_lds<T, NT>(GL * N, tid, b1, [&](int idx, int& dst) -> cx<T> {
      int m, g;
      fdivmod(idx, GL, rgl, m, g);
      dst = g * LS + lpad(m);
      return item_input<T>(I, G, N, NH, g, m);
    });
    __syncthreads();
  } else if (!fuse_in) {
#pragma unroll 1
    for (int idx = tid; idx < GL * N; idx += NT) {
      const int g = idx / N, m = idx - g * N;
      b1[g * LS + lpad(m)] = item_input<T>(I, G, N, NH, g, m);
    }
    __syncthreads();
  }
  auto natural = [&](int g, int m, const cx<T>& v, cx<T>* other) { other[g * LS + lpad(m)] = v; };
  // final values: straight to global memory (strided) or through LDS (rows)
  auto finish = [&](int g, int m, const cx<T>& v, cx<T>* other) {
    if (fuse_out)
      item_output<T>(I, G, g, m, v);
    else
      other[g * LS + lpad(m)] = v;
  };
  // the buffer the last pass's sink writes (the one it does not read): pass p
  // writes b0 when p is even, so the last pass of an odd-length plan writes b0
  constexpr bool last_in_b0 = (fft_npass(N) % 2 == 1);
  cx<T>* res = last_in_b0 ? b0 : b1;
  if (I.kind == FK_CONV || I.kind == FK_LINE) {
    // forward; W = conj(Z) * Bhat in natural order; forward again (from that buffer)
    auto sink_bhat = [&](int g, int m, const cx<T>& z, cx<T>* other) {
      const T b = conv_bl ? bl[m * GL + g] : item_bhat<T>(I, N, NH, g, m);
      other[g * LS + lpad(m)] = {z.x * b, -z.y * b};
    };
    CtRun<T, NT, N, 0, 1>::run(GL, rgl, tid, twl, b0, b1, false, fetch_b1, false, sink_bhat);
    __syncthreads();
    // W sits in res; the second transform's first pass reads it
    cx<T>* W = res;
    auto fetch_w = [&](int g, int m) -> cx<T> { return W[g * LS + lpad(m)]; };
    if (W == b0) {
      // the first pass must not write the buffer it reads: run with the roles swapped
      CtRun<T, NT, N, 0, 1>::run(GL, rgl, tid, twl, b1, b0, false, fetch_w, fuse_out, finish);
      res = last_in_b0 ? b1 : b0;
    } else {
      CtRun<T, NT, N, 0, 1>::run(GL, rgl, tid, twl, b0, b1, false, fetch_w, fuse_out, finish);
    }
  } else if (I.kind == FK_R2C) {
    CtRun<T, NT, N, 0, 1>::run(GL, rgl, tid, twl, b0, b1, false, fetch_b1, false, natural);
  } else {
    CtRun<T, NT, N, 0, 1>::run(GL, rgl, tid, twl, b0, b1, false, fetch_b1, fuse_out, finish);
  }
  if (I.kind == FK_R2C) {
    // Z = FFT(a + i b): A[k] = (Z[k] + conj(Z[-k])) / 2, B[k] = -i (Z[k] - conj(Z[-k])) / 2, k < nh
    __syncthreads();
    const int totp = G * NH;
#pragma unroll 1
    for (int idx = tid; idx < totp; idx += NT) {
      const int g = idx / NH, m = idx - g * NH;
      const cx<T>* Z = res + (g >> 1) * LS;
      const cx<T> z = Z[lpad(m)], zr = Z[lpad(m == 0 ? 0 : N - m)];
      const cx<T> o = (g & 1) ? cx<T>{T(0.5) * (z.y + zr.y), T(0.5) * (zr.x - z.x)}
                              : cx<T>{T(0.5) * (z.x + zr.x), T(0.5) * (z.y - zr.y)};
      gst<cx<T>>(reinterpret_cast<cx<T>*>(I.out) + (g * I.ls_out + (long)m * I.ps_out), o);
    }
  } else if (!fuse_out) {
    __syncthreads();
#pragma unroll 1
    for (int idx = tid; idx < GL * N; idx += NT) {
      const int g = idx / N, m = idx - g * N;
      item_output<T>(I, G, g, m, res[g * LS + lpad(m)]);
    }
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Run-time path: any length, every pass LDS to LDS (two buffers)
// ---------------------------------------------------------------------------
// LDS: the plan's twiddles (n slots) and roots (nroot slots), then two buffers of GL lines
template <typename T, int NT>
PF5D_DEV void fft_item_rt(const FftItem& I, const LinePlanDev& pl, const cx<T>* twl, const cx<T>* __restrict__ roots,
                          cx<T>* buf0, int tid) {
  const int n = pl.n, G = I.G, nh = n / 2 + 1, LS = line_stride(n);
  const bool paired = (I.kind == FK_R2C || I.kind == FK_C2R);
  const int GL = paired ? (G + 1) / 2 : G;
  cx<T>* buf1 = buf0 + GL * LS;
  const float rn = frcp(n), rgl = frcp(GL);
  const bool pm_in = I.ps_in != 1, pm_out = I.ps_out != 1;
  gather_lds<T, NT>(GL * n, tid, buf0, [&](int idx, int& dst) -> cx<T> {
    int g, m;
    if (pm_in)
      fdivmod(idx, GL, rgl, m, g);
    else
      fdivmod(idx, n, rn, g, m);
    dst = g * LS + lpad(m);
    return item_input<T>(I, G, n, nh, g, m);
  });
  __syncthreads();
  cx<T>* res = fft_lines<T, NT>(buf0, buf1, pl, LS, GL, twl, roots);
  if (I.kind == FK_CONV || I.kind == FK_LINE) {
    gather_lds<T, NT>(GL * n, tid, res, [&](int idx, int& dst) -> cx<T> {
      int g, m;
      fdivmod(idx, n, rn, g, m);
      dst = g * LS + lpad(m);
      const T b = item_bhat<T>(I, n, nh, g, m);
      const cx<T> z = res[dst];
      return {z.x * b, -z.y * b};
    });
    __syncthreads();
    res = fft_lines<T, NT>(res, (res == buf0) ? buf1 : buf0, pl, LS, GL, twl, roots);
  }
  if (I.kind == FK_R2C) {
    const int totp = G * nh;
    const float rnh = frcp(nh);
#pragma unroll 1
    for (int idx = tid; idx < totp; idx += NT) {
      int g, m;
      fdivmod(idx, nh, rnh, g, m);
      const cx<T>* Z = res + (g >> 1) * LS;
      const cx<T> z = Z[lpad(m)], zr = Z[lpad(m == 0 ? 0 : n - m)];
      const cx<T> o = (g & 1) ? cx<T>{T(0.5) * (z.y + zr.y), T(0.5) * (zr.x - z.x)}
                              : cx<T>{T(0.5) * (z.x + zr.x), T(0.5) * (z.y - zr.y)};
      gst<cx<T>>(reinterpret_cast<cx<T>*>(I.out) + (g * I.ls_out + (long)m * I.ps_out), o);
    }
  } else {
    const int totp = GL * n;
#pragma unroll 1
    for (int idx = tid; idx < totp; idx += NT) {
      int g, m;
      if (pm_out)
        fdivmod(idx, GL, rgl, m, g);
      else
        fdivmod(idx, n, rn, g, m);
      item_output<T>(I, G, g, m, res[g * LS + lpad(m)]);
    }
  }
  __syncthreads();
}

template <typename T, int NT>
__global__ void __launch_bounds__(NT) k_fft(const FftItem* __restrict__ items, int nitems,
                                            const LinePlanDev* __restrict__ plans, const cx<T>* __restrict__ tw,
                                            const cx<T>* __restrict__ roots, const SimState* __restrict__ st) {
  extern __shared__ __align__(16) unsigned char smem[];
  cx<T>* twl = reinterpret_cast<cx<T>*>(smem);
  int staged = -1;  // plan whose twiddles are in LDS
  // contiguous item ranges, consecutive ranges on one XCD: a stage's items for
  // one region of a simulation run on the XCD whose L2 holds what the previous
  // stage wrote there
  int ib, ie;
  block_range(nitems, ib, ie);
  for (int it = ib; it < ie; ++it) {
    const FftItem I = items[it];
    if (st && !(st[I.sim].status & ST_ACTIVE)) continue;
    const LinePlanDev& pl = plans[I.plan];
    // an empty asm keeps address arithmetic derived from the thread index from
    // being hoisted out of the item loop (it would hold registers across items)
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const bool tg = pl.tw_global != 0;
    if (I.plan != staged) {
      // the previous item's buffers are free: every item ends in a barrier.
      // LDS head: n - 1 twiddles, one spare slot, the generic radices' roots
      if (!tg) {
        for (int i = tid; i < pl.n - 1; i += NT) twl[i] = gld<cx<T>>(tw + pl.tw_off[0] + i);
        for (int i = tid; i < pl.nroot; i += NT) twl[pl.n + i] = gld<cx<T>>(roots + pl.root_off[0] + i);
      }
      staged = I.plan;
      __syncthreads();
    }
    cx<T>* buf = tg ? twl : twl + pl.n + pl.nroot;
    switch (pl.n) {  // the lengths of fft_ct_length
      case 64: fft_item_ct<T, NT, 64>(I, twl, buf, tid); break;
      case 96: fft_item_ct<T, NT, 96>(I, twl, buf, tid); break;
      case 128: fft_item_ct<T, NT, 128>(I, twl, buf, tid); break;
      case 192: fft_item_ct<T, NT, 192>(I, twl, buf, tid); break;
      case 256: fft_item_ct<T, NT, 256>(I, twl, buf, tid); break;
      case 512: fft_item_ct<T, NT, 512>(I, twl, buf, tid); break;
      default:
        if (tg)
          fft_item_rt<T, NT>(I, pl, tw + pl.tw_off[0], roots + pl.root_off[0], buf, tid);
        else
          fft_item_rt<T, NT>(I, pl, twl, twl + pl.n, buf, tid);
        break;
    }
  }
}

}  // namespace pf5d
