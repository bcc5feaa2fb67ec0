// Common device helpers for the phasefield5d MI355X step library.
//
// Build note: the whole library is compiled with -ffp-contract=off so that the
// FP64 kernels evaluate every product and sum separately, in the order of the
// reference Numba loops (fastmath=False).  FP32 kernels request fused
// multiply-adds explicitly through mad<float>.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PF5D_DEV __device__ __forceinline__
#define PF5D_HD __host__ __device__ __forceinline__

namespace pf5d {

// four-component composition / potential vector, 16 B (FP32) or 32 B (FP64)
template <typename T>
struct alignas(4 * sizeof(T)) V4 {
  T v[4];
};

// a*b + c: two roundings for double (reference order), one fused rounding for float
template <typename T>
PF5D_DEV T mad(T a, T b, T c);
template <>
PF5D_DEV double mad<double>(double a, double b, double c) { return a * b + c; }
template <>
PF5D_DEV float mad<float>(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

template <typename T>
PF5D_DEV V4<T> ld4(const T* p) { return *reinterpret_cast<const V4<T>*>(p); }
template <typename T>
PF5D_DEV void st4(T* p, const V4<T>& v) { *reinterpret_cast<V4<T>*>(p) = v; }

// Explicit address spaces: pointers that reach a kernel through a descriptor
// array are generic, and generic accesses compile to FLAT instructions, which
// wait on both the vector-memory and the LDS counters.  These helpers state
// the space so global data uses global_load/store and LDS data ds_read/write.
#define PF5D_GLOBAL __attribute__((address_space(1)))
#define PF5D_LDS __attribute__((address_space(3)))
// (class types cannot be copied out of a qualified address space, so the
// bytes move as a vector of 32-bit words; the copies fold away)
template <int NB>
struct Words {
  using type = unsigned int __attribute__((ext_vector_type(NB / 4)));
};
template <>
struct Words<4> {
  using type = unsigned int;
};
template <typename V>
PF5D_DEV V gld(const void* p) {
  using W = typename Words<sizeof(V)>::type;
  const W w = *(const PF5D_GLOBAL W*)p;
  V v;
  __builtin_memcpy(&v, &w, sizeof(V));
  return v;
}
template <typename V>
PF5D_DEV void gst(void* p, const V& v) {
  using W = typename Words<sizeof(V)>::type;
  W w;
  __builtin_memcpy(&w, &v, sizeof(V));
  *(PF5D_GLOBAL W*)p = w;
}
template <typename V>
PF5D_DEV V sld(const void* p) {
  using W = typename Words<sizeof(V)>::type;
  const W w = *(const PF5D_LDS W*)p;
  V v;
  __builtin_memcpy(&v, &w, sizeof(V));
  return v;
}
// streaming store: for outputs the next kernels read from other XCDs, so
// they need not stay dirty in this XCD's L2
template <typename V>
PF5D_DEV void gst_nt(void* p, const V& v) {
  using W = typename Words<sizeof(V)>::type;
  W w;
  __builtin_memcpy(&w, &v, sizeof(V));
  __builtin_nontemporal_store(w, (PF5D_GLOBAL W*)p);
}
// store of a stencil-stage output field: streaming in FP64, where an XCD's
// share of the field is about twice its L2 and nothing of it survives to the
// next kernel; a plain store in FP32, where the next kernel (same XCD, same
// slab) still finds part of it in L2
template <typename T, typename V>
PF5D_DEV void gst_field(void* p, const V& v) {
  if constexpr (sizeof(T) == 8)
    gst_nt<V>(p, v);
  else
    gst<V>(p, v);
}
template <typename V>
PF5D_DEV void sst(void* p, const V& v) {
  using W = typename Words<sizeof(V)>::type;
  W w;
  __builtin_memcpy(&w, &v, sizeof(V));
  *(PF5D_LDS W*)p = w;
}
// Accesses at a 32-bit byte offset from a base the whole wavefront shares
// (a simulation's field, in scalar registers): the compiler then issues
// global_load/store with the base as saddr and the offset in one VGPR, with no
// 64-bit address arithmetic per access.  Offsets stay below 4 GiB: the engine
// admits simulations of at most 2^32 / (4 sizeof(T)) cells.
template <typename V>
PF5D_DEV V gldo(const void* base, unsigned off) {
  return gld<V>(static_cast<const char*>(base) + off);
}
template <typename V>
PF5D_DEV void gsto(void* base, unsigned off, const V& v) {
  gst<V>(static_cast<char*>(base) + off, v);
}
template <typename T, typename V>
PF5D_DEV void gst_field_o(void* base, unsigned off, const V& v) {
  gst_field<T, V>(static_cast<char*>(base) + off, v);
}

// FP32 state remainder in 24 bits.  The FP64 state x is held as the nearest
// float hi and lo = rint((x - hi) / q), q = ulp(hi) / 2^23 (for zero and
// subnormal hi the ulp of the smallest normal float): |x - hi| <= ulp(hi) / 2,
// so |lo| <= 2^22, and hi + lo q is within q / 2 of x -- a 47-bit significand
// whose exact FP64 sum needs no rounding.  q is a power of two (2^(b - 173) for
// the biased exponent b of hi), so every scaling is exact.  Infinite or NaN hi
// keeps lo = 0.  A cell's four remainders pack into three 32-bit words (12 B
// against 16 for four float remainders).
struct alignas(4) Rem4 {
  unsigned w[3];
};
PF5D_HD int rem_exp(float hi) {
  unsigned bits;
  __builtin_memcpy(&bits, &hi, 4);
  const int b = (int)((bits >> 23) & 0xFFu);
  return b < 1 ? 1 : b;
}
PF5D_HD int rem_encode(double x, float hi) {
  const int b = rem_exp(hi);
  if (b == 255) return 0;
  return (int)__builtin_rint((x - (double)hi) * __builtin_ldexp(1.0, 173 - b));
}
PF5D_HD double rem_decode(float hi, int lo) { return (double)hi + (double)lo * __builtin_ldexp(1.0, rem_exp(hi) - 173); }
// four 24-bit two's-complement values: v0 in w0[0:24), v1 in w0[24:32) and
// w1[0:16), v2 in w1[16:32) and w2[0:8), v3 in w2[8:32)
PF5D_HD Rem4 rem_pack(const int v[4]) {
  Rem4 r;
  r.w[0] = ((unsigned)v[0] & 0xFFFFFFu) | ((unsigned)v[1] << 24);
  r.w[1] = (((unsigned)v[1] >> 8) & 0xFFFFu) | ((unsigned)v[2] << 16);
  r.w[2] = (((unsigned)v[2] >> 16) & 0xFFu) | ((unsigned)v[3] << 8);
  return r;
}
PF5D_HD void rem_unpack(const Rem4& r, int v[4]) {
  v[0] = (int)(r.w[0] << 8) >> 8;
  v[1] = (int)(((r.w[0] >> 24) | (r.w[1] << 8)) << 8) >> 8;
  v[2] = (int)(((r.w[1] >> 16) | (r.w[2] << 16)) << 8) >> 8;
  v[3] = (int)r.w[2] >> 8;
}
// a cell's remainders in HBM, as three 32-bit accesses (a 12-byte vector type
// would be padded to 16 bytes and read past the last cell)
PF5D_DEV Rem4 ld_rem(const Rem4* p) {
  const unsigned* q = reinterpret_cast<const unsigned*>(p);
  Rem4 r;
  r.w[0] = gld<unsigned>(q);
  r.w[1] = gld<unsigned>(q + 1);
  r.w[2] = gld<unsigned>(q + 2);
  return r;
}
PF5D_DEV void st_rem(Rem4* p, const Rem4& r) {
  unsigned* q = reinterpret_cast<unsigned*>(p);
  gst<unsigned>(q, r.w[0]);
  gst<unsigned>(q + 1, r.w[1]);
  gst<unsigned>(q + 2, r.w[2]);
}

template <typename T>
PF5D_DEV V4<T> ld4g(const T* p) { return gld<V4<T>>(p); }
template <typename T>
PF5D_DEV void st4g(T* p, const V4<T>& v) { gst<V4<T>>(p, v); }

PF5D_HD int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// float reciprocal of a positive integer for fdivmod: the hardware estimate
// (1 ulp) is enough, fdivmod corrects the quotient
PF5D_DEV float frcp(int n) { return __builtin_amdgcn_rcpf((float)n); }

// q = idx / cols, r = idx % cols for idx < 2^20, cols < 2^11, via a float
// reciprocal (exact or frcp) and one correction step (avoids the
// integer-division sequence): the quotient estimate is off by at most one
PF5D_DEV void fdivmod(int idx, int cols, float rcp, int& q, int& r) {
  q = (int)((float)idx * rcp);
  r = idx - q * cols;
  if (r < 0) {
    q -= 1;
    r += cols;
  } else if (r >= cols) {
    q += 1;
    r -= cols;
  }
}

// non-negative IEEE values order like their bit patterns: max via integer
// atomics (device scope); returns the previous bits
PF5D_DEV unsigned long long atomic_max_nonneg(double* addr, double v) {
  return atomicMax(reinterpret_cast<unsigned long long*>(addr), (unsigned long long)__double_as_longlong(v));
}
PF5D_DEV unsigned long long atomic_max_nonneg(float* addr, float v) {
  return atomicMax(reinterpret_cast<unsigned int*>(addr), (unsigned int)__float_as_uint(v));
}

// max that ignores NaN (the reference abs_max only replaces on v > m)
template <typename T>
PF5D_DEV T max_ignore_nan(T m, T v) { return (v > m) ? v : m; }

// wavefront (64 lanes) max reduction by butterfly shuffles
template <typename T>
PF5D_DEV T wave_max(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    T o = __shfl_xor(v, off, 64);
    v = (o > v) ? o : v;
  }
  return v;
}
PF5D_DEV double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// XCDs (dies with their own L2) of the chip; workgroups are handed to them
// round robin, workgroup b to XCD b % NXCD
constexpr int NXCD = 8;

// This block's contiguous item range [ib, ie) (the items split as evenly as
// possible).  With a grid that is a multiple of NXCD, XCD j takes the j-th
// eighth of the list: work lists ordered by region (planes of a simulation)
// keep each region on one XCD from kernel to kernel, so what one kernel wrote
// is in the L2 the next one reads, and a block's neighbouring planes were
// brought in by the blocks working next to it.
PF5D_DEV void block_range(int nitems, int& ib, int& ie) {
  const int G = gridDim.x, b = blockIdx.x;
  const int L = (G % NXCD == 0) ? (b % NXCD) * (G / NXCD) + b / NXCD : b;
  ib = (int)((long)L * nitems / G);
  ie = (int)((long)(L + 1) * nitems / G);
}

enum StatusBits { ST_ACTIVE = 1, ST_HARD_OOB = 2, ST_SOFT_OOB = 4, ST_DONE = 8 };

// per-simulation mutable state (FP64 scalars on the device)
struct SimState {
  double dt, t, last_xmax;
  double mass[4], mean[4];
  double smax, smin, cmin, cmax, nonfinite;
  long long steps;
  int status;
  int pad;
};

}  // namespace pf5d
