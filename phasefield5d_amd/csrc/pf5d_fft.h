// Batched mixed-radix FFT in LDS for the rank-1 elastic term
//     psi = irfftn( Bhat * rfftn(s) )     (phasefield5d/solver/elastic.py:246-263)
//
// Algorithm: Stockham autosort FFT in the formulation of Govindaraju, Lloyd,
// Dotsenko, Smith, Manferdelli, "High performance discrete Fourier transforms
// on graphics processors", SC08: for each pass with radix R and sub-transform
// size Ns, butterfly j reads x[j + r n/R], twiddles by exp(-2 pi i r (j mod Ns)
// / (Ns R)), applies a radix-R DFT and writes y[(j/Ns) Ns R + (j mod Ns) + r Ns].
// Radices 2, 3, 4, 5, 8 are hand-written butterflies; any other factor uses a
// direct radix-R DFT from a per-plan root table.  Twiddles are computed on the
// host in extended precision and rounded to the working type.
//
// Every work item transforms G lines of length n.  Point m of line g sits at
// g*ls + m*ps (ps = 1 for rows along the contiguous axis, ls = 1 for pencils
// along a strided axis, so global loads are coalesced either way).
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
PF5D_DEV cx<T> cmul(cx<T> a, cx<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
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

// LDS position of point e of a line: one spare slot after every 8 points, so
// the power-of-two strides of the Stockham reads and writes spread over the
// banks; a line of n points takes line_stride(n) slots
PF5D_DEV constexpr int lpad(int e) { return e + (e >> 3); }
PF5D_HD constexpr int line_stride(int n) { return n + (n >> 3) + 1; }

// one Stockham pass over G lines (line stride LS): src -> dst.  tw holds this
// pass's twiddles (in LDS), roots the generic radix's roots of unity.
template <typename T, int NT>
PF5D_DEV void stockham_pass(const cx<T>* __restrict__ src, cx<T>* __restrict__ dst, int n, int LS, int G, int R,
                            int Ns, const cx<T>* __restrict__ tw, const cx<T>* __restrict__ roots) {
  const int nb = n / R;
  const int tot = G * nb;
  const float rnb = 1.0f / (float)nb, rns = 1.0f / (float)Ns;
  for (int b = threadIdx.x; b < tot; b += NT) {
    int g, j, q, k;
    fdivmod(b, nb, rnb, g, j);
    fdivmod(j, Ns, rns, q, k);
    const cx<T>* s = src + g * LS;
    cx<T>* d = dst + g * LS;
    const int ob = q * Ns * R + k;
    const cx<T>* w = tw + k * (R - 1);  // w[r-1] is the twiddle of input r >= 1
    if (R == 2 || R == 3 || R == 4 || R == 5 || R == 8) {
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
    } else {
      // direct radix-R DFT: y_q = sum_p (x_p w_p) root^(p q mod R)
      for (int qq = 0; qq < R; ++qq) {
        cx<T> acc = s[lpad(j)];
        int e = 0;
        for (int p = 1; p < R; ++p) {
          e += qq;
          if (e >= R) e -= R;
          acc = cadd(acc, cmul(cmul(s[lpad(j + p * nb)], w[p - 1]), gld<cx<T>>(roots + e)));
        }
        d[lpad(ob + qq * Ns)] = acc;
      }
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
    stockham_pass<T, NT>(a, b, pl.n, LS, G, R, Ns, tw + (pl.tw_off[p] - pl.tw_off[0]), roots + pl.root_off[p]);
    __syncthreads();
    cx<T>* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

// ---------------------------------------------------------------------------
// Compile-time plans for the line lengths of the usual grids: the radix list
// (the host's factorisation order: 8s, 4s, 2s, then 3, 5, 7) is fixed at
// compile time, so every index computation of a pass folds to shifts and
// constant multiplies and the butterfly of each pass is a straight-line block.
// ---------------------------------------------------------------------------
PF5D_HD constexpr int ct_radix(int n, int p) {
  int m = n, i = 0;
  const int order[6] = {8, 4, 2, 3, 5, 7};
  for (int q = 0; q < 6; ++q) {
    const int R = order[q];
    while (m % R == 0 && m >= R) {
      if (i == p) return R;
      m /= R;
      ++i;
    }
  }
  return 0;
}
PF5D_HD constexpr int ct_npass(int n) {
  int p = 0;
  while (ct_radix(n, p) != 0) ++p;
  return p;
}
// lengths with a compile-time plan
PF5D_HD constexpr bool ct_length(int n) {
  return n == 64 || n == 96 || n == 128 || n == 160 || n == 192 || n == 256 || n == 384 || n == 512;
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
  else
    dft8(v);
}

// one pass with radix R after sub-transforms of size NS, lines of N points
template <typename T, int NT, int N, int R, int NS>
PF5D_DEV void stockham_ct(const cx<T>* __restrict__ src, cx<T>* __restrict__ dst, int G,
                          const cx<T>* __restrict__ tw, int tid) {
  constexpr int NB = N / R, LS = line_stride(N);
  const int tot = G * NB;
#pragma unroll 1
  for (int b = tid; b < tot; b += NT) {
    const int g = b / NB, j = b - g * NB;
    const int k = j % NS, q = j / NS;
    const cx<T>* s = src + g * LS;
    cx<T>* d = dst + g * LS + 0;
    cx<T> v[R];
    v[0] = s[lpad(j)];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      const cx<T> x = s[lpad(j + r * NB)];
      if constexpr (NS == 1)
        v[r] = x;
      else
        v[r] = cmul(x, tw[k * (R - 1) + r - 1]);
    }
    dft_r<R>(v);
    const int ob = q * NS * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[lpad(ob + r * NS)] = v[r];
  }
}

// the passes of the compile-time plan; tw holds this pass's twiddles first
// (pass p's twiddles start at NS_p - 1 in the plan's table)
template <typename T, int NT, int N, int P, int NS>
struct CtPasses {
  static PF5D_DEV cx<T>* run(cx<T>* a, cx<T>* b, int G, const cx<T>* tw, int tid) {
    if constexpr (P == ct_npass(N)) {
      return a;
    } else {
      constexpr int R = ct_radix(N, P);
      stockham_ct<T, NT, N, R, NS>(a, b, G, tw, tid);
      __syncthreads();
      return CtPasses<T, NT, N, P + 1, NS * R>::run(b, a, G, tw + NS * (R - 1), tid);
    }
  }
};

// q = idx / n, r = idx % n with n = N when N > 0 (compile time), else runtime
template <int N>
PF5D_DEV void divmod_n(int idx, int n, float rn, int& q, int& r) {
  if constexpr (N > 0) {
    q = idx / N;
    r = idx - q * N;
  } else {
    fdivmod(idx, n, rn, q, r);
  }
}

// Work item layout in LDS: the plan's n - 1 twiddles (n slots), then two line
// buffers of GL lines (stride line_stride(n)).  R2C and C2R items carry their
// G real lines as GL = ceil(G/2) complex lines (line 2q in the real part, line
// 2q+1 in the imaginary part), which halves their transform work.
// N > 0: compile-time length N; N = 0: the plan's length at run time.
template <typename T, int NT, int N>
PF5D_DEV void fft_item(const FftItem& I, const LinePlanDev& pl, const cx<T>* __restrict__ tw,
                       const cx<T>* __restrict__ roots, unsigned char* smem) {
  // the thread index is passed through an empty asm statement, so that address
  // arithmetic derived from it is not hoisted out of the caller's item loop
  // (holding it across items would cost more registers than recomputing it)
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int n = N > 0 ? N : pl.n;
  const int G = I.G, nh = n / 2 + 1, LS = line_stride(n);
  const bool paired = (I.kind == FK_R2C || I.kind == FK_C2R);
  const int GL = paired ? (G + 1) / 2 : G;
  cx<T>* twl = reinterpret_cast<cx<T>*>(smem);
  cx<T>* buf0 = twl + n;
  cx<T>* buf1 = buf0 + GL * LS;
  const float rn = 1.0f / (float)n, rg = 1.0f / (float)G;
  // ---- load
  if (paired) {
    // rows along the contiguous axis (point stride 1); pair q = lines 2q, 2q+1
    const int totq = GL * n;
    #pragma unroll 1
    for (int idx = tid; idx < totq; idx += NT) {
      int q, m;
      divmod_n<N>(idx, n, rn, q, m);
      const int g0 = 2 * q, g1 = 2 * q + 1;
      const bool has1 = g1 < G;
      cx<T> v;
      if (I.kind == FK_R2C) {
        const T* in = reinterpret_cast<const T*>(I.in);
        v.x = gld<T>(in + (g0 * I.ls_in + (long)m * I.ps_in));
        v.y = has1 ? gld<T>(in + (g1 * I.ls_in + (long)m * I.ps_in)) : T(0);
      } else {
        // Hermitian extension of a half spectrum, conjugated for the inverse:
        // conj(A_full[m]) = conj(A[m]) (m < nh) or A[n - m]; then conj(A + i B) = conj(A) - i conj(B)
        const cx<T>* C = reinterpret_cast<const cx<T>*>(I.in);
        auto cfull = [&](int g) -> cx<T> {
          const cx<T>* L = C + g * I.ls_in;
          return (m < nh) ? conjg(gld<cx<T>>(L + (long)m * I.ps_in)) : gld<cx<T>>(L + (long)(n - m) * I.ps_in);
        };
        const cx<T> ca = cfull(g0);
        const cx<T> cb = has1 ? cfull(g1) : cx<T>{T(0), T(0)};
        v = {ca.x + cb.y, ca.y - cb.x};
      }
      buf0[q * LS + lpad(m)] = v;
    }
  } else {
    // point-major order for strided lines: consecutive threads take consecutive lines
    const bool rowmajor = (I.ps_in == 1);
    const int tot = G * n;
    #pragma unroll 1
    for (int idx = tid; idx < tot; idx += NT) {
      int g, m;
      if (rowmajor)
        divmod_n<N>(idx, n, rn, g, m);
      else
        fdivmod(idx, G, rg, m, g);
      cx<T> v;
      if (I.kind == FK_LINE) {
        v = {gld<T>(reinterpret_cast<const T*>(I.in) + (g * I.ls_in + (long)m * I.ps_in)), T(0)};
      } else {
        v = gld<cx<T>>(reinterpret_cast<const cx<T>*>(I.in) + (g * I.ls_in + (long)m * I.ps_in));
        if (I.kind == FK_INV) v = conjg(v);
      }
      buf0[g * LS + lpad(m)] = v;
    }
  }
  __syncthreads();
  auto transform = [&](cx<T>* a, cx<T>* b) -> cx<T>* {
    if constexpr (N > 0)
      return CtPasses<T, NT, N, 0, 1>::run(a, b, GL, twl, tid);
    else
      return fft_lines<T, NT>(a, b, pl, LS, GL, twl, roots);
  };
  cx<T>* res = transform(buf0, buf1);
  if (I.kind == FK_CONV || I.kind == FK_LINE) {
    // W = conj(Z) * Bhat, then a second forward transform; result = conj(V)
    const int tl = GL * n;
    #pragma unroll 1
    for (int idx = tid; idx < tl; idx += NT) {
      int g, m;
      divmod_n<N>(idx, n, rn, g, m);
      T b;
      if (I.kind == FK_LINE)
        b = gld<T>(reinterpret_cast<const T*>(I.bhat) + ((m < nh) ? m : n - m));
      else
        b = gld<T>(reinterpret_cast<const T*>(I.bhat) + (g * I.ls_b + (long)m * I.ps_b));
      cx<T>& z = res[g * LS + lpad(m)];
      z = {z.x * b, -z.y * b};
    }
    __syncthreads();
    res = transform(res, (res == buf0) ? buf1 : buf0);
  }
  // ---- store
  if (paired) {
    if (I.kind == FK_R2C) {
      const int totp = G * nh;
      const float rnh = 1.0f / (float)nh;
      #pragma unroll 1
      for (int idx = tid; idx < totp; idx += NT) {
        int g, m;
        divmod_n<(N > 0 ? N / 2 + 1 : 0)>(idx, nh, rnh, g, m);
        // Z = FFT(a + i b): A[k] = (Z[k] + conj(Z[-k])) / 2, B[k] = -i (Z[k] - conj(Z[-k])) / 2
        const cx<T>* Z = res + (g >> 1) * LS;
        const cx<T> z = Z[lpad(m)], zr = Z[lpad(m == 0 ? 0 : n - m)];
        const cx<T> o = (g & 1) ? cx<T>{T(0.5) * (z.y + zr.y), T(0.5) * (zr.x - z.x)}
                                : cx<T>{T(0.5) * (z.x + zr.x), T(0.5) * (z.y - zr.y)};
        gst<cx<T>>(reinterpret_cast<cx<T>*>(I.out) + (g * I.ls_out + (long)m * I.ps_out), o);
      }
    } else {
      const int totp = G * n;
      #pragma unroll 1
      for (int idx = tid; idx < totp; idx += NT) {
        int g, m;
        divmod_n<N>(idx, n, rn, g, m);
        // z = conj(W): line 2q is Re W, line 2q+1 is -Im W
        const cx<T> w = res[(g >> 1) * LS + lpad(m)];
        gst<T>(reinterpret_cast<T*>(I.out) + (g * I.ls_out + (long)m * I.ps_out), (g & 1) ? -w.y : w.x);
      }
    }
  } else {
    const bool rowmajor_o = (I.ps_out == 1);
    const int totp = G * n;
    #pragma unroll 1
    for (int idx = tid; idx < totp; idx += NT) {
      int g, m;
      if (rowmajor_o)
        divmod_n<N>(idx, n, rn, g, m);
      else
        fdivmod(idx, G, rg, m, g);
      const cx<T> v = res[g * LS + lpad(m)];
      if (I.kind == FK_LINE) {
        gst<T>(reinterpret_cast<T*>(I.out) + (g * I.ls_out + (long)m * I.ps_out), v.x);
      } else {
        const cx<T> o = (I.kind == FK_INV || I.kind == FK_CONV) ? conjg(v) : v;
        gst<cx<T>>(reinterpret_cast<cx<T>*>(I.out) + (g * I.ls_out + (long)m * I.ps_out), o);
      }
    }
  }
  __syncthreads();
}

template <typename T, int NT>
__global__ void __launch_bounds__(NT) k_fft(const FftItem* __restrict__ items, int nitems,
                                            const LinePlanDev* __restrict__ plans, const cx<T>* __restrict__ tw,
                                            const cx<T>* __restrict__ roots, const SimState* __restrict__ st) {
  extern __shared__ __align__(16) unsigned char smem[];
  int staged = -1;  // plan whose twiddles are in LDS
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const FftItem I = items[it];
    if (st && !(st[I.sim].status & ST_ACTIVE)) continue;
    const LinePlanDev& pl = plans[I.plan];
    if (I.plan != staged) {
      // the previous item's buffers are free: its store loop ended in a barrier
      cx<T>* twl = reinterpret_cast<cx<T>*>(smem);
      for (int i = threadIdx.x; i < pl.n - 1; i += NT) twl[i] = gld<cx<T>>(tw + pl.tw_off[0] + i);
      staged = I.plan;
    }
    switch (pl.n) {
      case 64: fft_item<T, NT, 64>(I, pl, tw, roots, smem); break;
      case 96: fft_item<T, NT, 96>(I, pl, tw, roots, smem); break;
      case 128: fft_item<T, NT, 128>(I, pl, tw, roots, smem); break;
      case 160: fft_item<T, NT, 160>(I, pl, tw, roots, smem); break;
      case 192: fft_item<T, NT, 192>(I, pl, tw, roots, smem); break;
      case 256: fft_item<T, NT, 256>(I, pl, tw, roots, smem); break;
      case 384: fft_item<T, NT, 384>(I, pl, tw, roots, smem); break;
      case 512: fft_item<T, NT, 512>(I, pl, tw, roots, smem); break;
      default: fft_item<T, NT, 0>(I, pl, tw, roots, smem); break;
    }
  }
}

}  // namespace pf5d
