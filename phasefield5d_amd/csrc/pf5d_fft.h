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

// one Stockham pass over G lines: src -> dst
template <typename T, int NT>
PF5D_DEV void stockham_pass(const cx<T>* __restrict__ src, cx<T>* __restrict__ dst, int n, int G, int R, int Ns,
                            const cx<T>* __restrict__ tw, const cx<T>* __restrict__ roots) {
  const int nb = n / R;
  const int tot = G * nb;
  for (int b = threadIdx.x; b < tot; b += NT) {
    const int g = b / nb, j = b - g * nb;
    const int k = j % Ns;
    const cx<T>* s = src + g * n + j;
    cx<T>* d = dst + g * n + (j / Ns) * Ns * R + k;
    const cx<T>* w = tw + k * (R - 1);  // w[r-1] is the twiddle of input r >= 1
    if (R == 2 || R == 3 || R == 4 || R == 5 || R == 8) {
      cx<T> v[8];
      v[0] = s[0];
#pragma unroll
      for (int r = 1; r < 8; ++r)
        if (r < R) v[r] = cmul(s[r * nb], w[r - 1]);
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
        if (r < R) d[r * Ns] = v[r];
    } else {
      // direct radix-R DFT: y_q = sum_p (x_p w_p) root^(p q mod R)
      for (int q = 0; q < R; ++q) {
        cx<T> acc = s[0];
        int e = 0;
        for (int p = 1; p < R; ++p) {
          e += q;
          if (e >= R) e -= R;
          acc = cadd(acc, cmul(cmul(s[p * nb], w[p - 1]), roots[e]));
        }
        d[q * Ns] = acc;
      }
    }
  }
}

// full forward FFT of G lines held in buf[0]; returns the buffer holding the result
template <typename T, int NT>
PF5D_DEV cx<T>* fft_lines(cx<T>* buf0, cx<T>* buf1, const LinePlanDev& pl, int G, const cx<T>* tw,
                          const cx<T>* roots) {
  cx<T>* a = buf0;
  cx<T>* b = buf1;
  int Ns = 1;
  for (int p = 0; p < pl.npass; ++p) {
    const int R = pl.radix[p];
    stockham_pass<T, NT>(a, b, pl.n, G, R, Ns, tw + pl.tw_off[p], roots + pl.root_off[p]);
    __syncthreads();
    cx<T>* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}

template <typename T, int NT>
__global__ void __launch_bounds__(NT) k_fft(const FftItem* __restrict__ items, int nitems,
                                            const LinePlanDev* __restrict__ plans, const cx<T>* __restrict__ tw,
                                            const cx<T>* __restrict__ roots, const SimState* __restrict__ st) {
  extern __shared__ __align__(16) unsigned char smem[];
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {
    const FftItem I = items[it];
    if (st && !(st[I.sim].status & ST_ACTIVE)) continue;
    const LinePlanDev& pl = plans[I.plan];
    const int n = pl.n, G = I.G, nh = n / 2 + 1;
    cx<T>* buf0 = reinterpret_cast<cx<T>*>(smem);
    cx<T>* buf1 = buf0 + (long)G * n;
    // ---- load
    const int tot = G * n;
    const bool rowmajor = (I.ps_in == 1);
    for (int idx = threadIdx.x; idx < tot; idx += NT) {
      int g, m;
      if (rowmajor) {
        g = idx / n;
        m = idx - g * n;
      } else {
        m = idx / G;
        g = idx - m * G;
      }
      cx<T> v;
      if (I.kind == FK_R2C || I.kind == FK_LINE) {
        v = {gld<T>(reinterpret_cast<const T*>(I.in) + (g * I.ls_in + (long)m * I.ps_in)), T(0)};
      } else if (I.kind == FK_C2R) {
        const cx<T>* C = reinterpret_cast<const cx<T>*>(I.in) + g * I.ls_in;
        v = (m < nh) ? conjg(gld<cx<T>>(C + (long)m * I.ps_in)) : gld<cx<T>>(C + (long)(n - m) * I.ps_in);
      } else {
        v = gld<cx<T>>(reinterpret_cast<const cx<T>*>(I.in) + (g * I.ls_in + (long)m * I.ps_in));
        if (I.kind == FK_INV) v = conjg(v);
      }
      buf0[g * n + m] = v;
    }
    __syncthreads();
    cx<T>* res = fft_lines<T, NT>(buf0, buf1, pl, G, tw, roots);
    if (I.kind == FK_CONV || I.kind == FK_LINE) {
      // W = conj(Z) * Bhat, then a second forward transform; result = conj(V)
      for (int idx = threadIdx.x; idx < tot; idx += NT) {
        const int g = idx / n, m = idx - g * n;
        T b;
        if (I.kind == FK_LINE)
          b = gld<T>(reinterpret_cast<const T*>(I.bhat) + ((m < nh) ? m : n - m));
        else
          b = gld<T>(reinterpret_cast<const T*>(I.bhat) + (g * I.ls_b + (long)m * I.ps_b));
        const cx<T> z = res[idx];
        res[idx] = {z.x * b, -z.y * b};
      }
      __syncthreads();
      cx<T>* other = (res == buf0) ? buf1 : buf0;
      res = fft_lines<T, NT>(res, other, pl, G, tw, roots);
    }
    // ---- store
    const bool rowmajor_o = (I.ps_out == 1);
    const int mlim = (I.kind == FK_R2C) ? nh : n;
    const int totp = G * mlim;
    for (int idx = threadIdx.x; idx < totp; idx += NT) {
      int g, m;
      if (rowmajor_o) {
        g = idx / mlim;
        m = idx - g * mlim;
      } else {
        m = idx / G;
        g = idx - m * G;
      }
      const cx<T> v = res[g * n + m];
      if (I.kind == FK_C2R || I.kind == FK_LINE) {
        gst<T>(reinterpret_cast<T*>(I.out) + (g * I.ls_out + (long)m * I.ps_out), v.x);
      } else {
        cx<T> o = v;
        if (I.kind == FK_INV || I.kind == FK_CONV) o = conjg(o);
        gst<cx<T>>(reinterpret_cast<cx<T>*>(I.out) + (g * I.ls_out + (long)m * I.ps_out), o);
      }
    }
    __syncthreads();
  }
}

}  // namespace pf5d
