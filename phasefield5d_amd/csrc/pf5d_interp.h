// 4D CALPHAD multilinear interpolation of the nine properties
// [mu_Mn, mu_Ni, mu_Co, mu_Cu, Mob_Fe, Mob_Mn, Mob_Ni, Mob_Co, Mob_Cu].
//
// Follows phasefield5d/solver/interpolation.py:67-147 and :154-211:
//   * linear region: sum(X) <= 1 - 4h - 1e-9 (interpolation.py:195);
//   * per axis u = X/h; u <= 0 -> (i=0, t=0); u >= num-1 -> (i=num-2, t=1);
//     else i = int(u), t = u - i (interpolation.py:110-121);
//   * 16 weights prod(t or 1-t), summed c0000..c1111 with i3 fastest
//     (interpolation.py:126-147), left to right;
//   * outside the linear region the value of the nearest table node in
//     Euclidean X (the reference's cKDTree query, interpolation.py:203-209).
//     For a table whose rows are the simplex grid nodes, the nearest node to a
//     point with X >= 0, sum(X) <= 1 lies within [floor(u)-1, floor(u)+2] on
//     every axis (the floor node is closer than 2h), so that 4^4 window is
//     searched exhaustively; ties keep the lowest row-major node.
//
// Device table layout: dense (num^4, NPAD) with the nine properties first and
// NPAD = 10 (FP64, five 16-byte loads per corner) or 12 (FP32, three 16-byte
// loads per corner).  Nodes outside the simplex hold NaN, as in build_4d_grid.
#pragma once
#include "pf5d_common.h"

namespace pf5d {

template <typename T>
struct TablePad;
template <>
struct TablePad<double> { static constexpr int value = 10; };
template <>
struct TablePad<float> { static constexpr int value = 12; };

template <typename T>
struct TableView {
  const T* data;   // (num^4, NPAD)
  int num;
  T inv_step;      // 1/h (host: 1.0/h in FP64, rounded for FP32)
  T thresh;        // 1 - 4h - 1e-9
  double step;     // h, for node coordinates in the nearest search
  const T* cubes;  // cube table (below), or null
  int cube_s;      // cubes with lower-corner index sum <= cube_s are in it
};

// nearest table row for points past the linear region (defined below)
template <typename T, bool WIDE>
PF5D_DEV void interp_nearest(const TableView<T>& tv, const T X[4], T P[9]);

// Cube table: for every hypercube of the linear region (lower corner i with
// i0+i1+i2+i3 <= S, S = num - 5) the 16 corners' nine properties, stored in
// the order the interpolation reads them, CUBE_VALUES values per cube: for
// each 16-byte group g of properties, the 16 corners (c0000 .. c1111, i3
// fastest) one after the other, then the ninth property of the 16 corners.  A
// point's gather then reads 576 contiguous bytes (FP32) instead of 16 table
// rows scattered over the 4D grid.  Cubes are ranked lexicographically among
// the tuples with sum <= S.
constexpr int CUBE_VALUES = 144;
PF5D_HD long simplex_rank4(int i0, int i1, int i2, int i3, int S) {
  // number of k-tuples of non-negative integers with sum <= r: C(r + k, k)
  auto f2 = [](long r) { return (r + 1) * (r + 2) / 2; };
  auto f3 = [](long r) { return (r + 1) * (r + 2) * (r + 3) / 6; };
  auto f4 = [](long r) { return (r + 1) * (r + 2) * (r + 3) * (r + 4) / 24; };
  const long R1 = S - i0, R2 = R1 - i1;
  return (f4(S) - f4(R1)) + (f3(R1) - f3(R1 - i1)) + (f2(R2) - f2(R2 - i2)) + i3;
}
// simplex_rank4 in 32-bit arithmetic, for S <= CUBE_MAX_S: the binomials
// C(r + k, k) are built as C(r+2,2) = (r+1)(r+2)/2, C(r+3,3) = C(r+2,2)(r+3)/3,
// C(r+4,4) = C(r+3,3)(r+4)/4, and for r <= 390 every factor is below 2^24 and
// every product below 2^32, so each multiply is a 24-bit one (the masks state
// the range to the compiler and change no value) and each division exact.  The
// interpolation takes the rank per point; 64-bit products and divisions cost it
// several times the instructions.
constexpr int CUBE_MAX_S = 390;
PF5D_DEV unsigned binom2(unsigned r) { return ((r & 0xFFFFu) + 1) * ((r & 0xFFFFu) + 2) >> 1; }
PF5D_DEV unsigned binom3(unsigned r) { return (binom2(r) & 0xFFFFFFu) * ((r & 0xFFFFu) + 3) / 3; }
PF5D_DEV unsigned binom4(unsigned r) { return (binom3(r) & 0xFFFFFFu) * ((r & 0xFFFFu) + 4) >> 2; }
PF5D_DEV unsigned cube_rank(int i0, int i1, int i2, int i3, int S) {
  const unsigned R1 = S - i0, R2 = R1 - i1;
  return (binom4(S) - binom4(R1)) + (binom3(R1) - binom3(R1 - i1)) + (binom2(R2) - binom2(R2 - i2)) + i3;
}
PF5D_HD long simplex_count4(int S) { return S < 0 ? 0 : (long)(S + 1) * (S + 2) * (S + 3) * (S + 4) / 24; }

template <typename T>
PF5D_DEV void load_corner(const T* p, T g[TablePad<T>::value]);

template <>
PF5D_DEV void load_corner<double>(const double* p, double g[10]) {
  const double2* q = reinterpret_cast<const double2*>(p);
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    double2 v = q[i];
    g[2 * i] = v.x;
    g[2 * i + 1] = v.y;
  }
}
template <>
PF5D_DEV void load_corner<float>(const float* p, float g[12]) {
  const float4* q = reinterpret_cast<const float4*>(p);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    float4 v = q[i];
    g[4 * i] = v.x;
    g[4 * i + 1] = v.y;
    g[4 * i + 2] = v.z;
    g[4 * i + 3] = v.w;
  }
}

// interpolation.py:110-121: u <= 0 -> (0, 0); u >= num-1 -> (num-2, 1); else
// (int(u), u - int(u)).  Branch-free: the floor clamped to [0, num-2], then
// t = u - i clamped to [0, 1] (+0 turns a -0 into +0 as the first case has it).
// In the open interval both give floor(u) and the same subtraction; below and
// above, the clamps give exactly 0 and 1.
template <typename T>
PF5D_DEV void axis_index(T u, int num, int& i, T& t) {
  const T fi = fmin(fmax(floor(u), T(0)), T(num - 2));
  i = (int)fi;
  t = fmin(fmax(u - fi, T(0)), T(1)) + T(0);
}

// 16-byte vector of the table element type
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  using type = float4;
  static constexpr int n = 4;
};
template <>
struct Vec16<double> {
  using type = double2;
  static constexpr int n = 2;
};
template <typename V>
PF5D_DEV auto vget(const V& v, int i) {
  if constexpr (sizeof(v.x) == 4) {
    return i == 0 ? v.x : i == 1 ? v.y : i == 2 ? v.z : v.w;
  } else {
    return i == 0 ? v.x : v.y;
  }
}

// The 16 corner rows are gathered in batches of one 16-byte column per corner
// (16 independent loads in flight per batch from LDS, 8 from HBM; one wait per
// batch); within each property the sum still runs c0000 .. c1111 left to right.  ``base`` points at
// the row of the lower corner; n1, n2, n3 are the element strides of axes 2, 1, 0
// (the stride of axis 3 is one row).  The rows may live in the HBM table or in
// an LDS copy of a box of nodes: the arithmetic is the same.
// the 16 corner weights ((f0 f1) f2) f3 as a product tree: same roundings as the
// reference's left-to-right product (interpolation.py:126-134), 28 multiplies
template <typename T>
PF5D_DEV void corner_weights(T t0, T t1, T t2, T t3, T w[16]) {
  const T s0 = T(1) - t0, s1 = T(1) - t1, s2 = T(1) - t2, s3 = T(1) - t3;
  T w01[4], w012[8];
#pragma unroll
  for (int a = 0; a < 4; ++a) w01[a] = ((a >> 1) ? t0 : s0) * ((a & 1) ? t1 : s1);
#pragma unroll
  for (int a = 0; a < 8; ++a) w012[a] = w01[a >> 1] * ((a & 1) ? t2 : s2);
#pragma unroll
  for (int b = 0; b < 16; ++b) w[b] = w012[b >> 1] * ((b & 1) ? t3 : s3);
}

template <typename T, bool FROM_LDS = false>
PF5D_DEV void interp_core(const T* base, int n1, int n2, int n3, T t0, T t1, T t2, T t3, T P[9]) {
  constexpr int NP = TablePad<T>::value;
  using VT = typename Vec16<T>::type;
  constexpr int VN = Vec16<T>::n;
  constexpr int NB = (9 + VN - 1) / VN;  // batches: 3 (FP32) or 5 (FP64)
  // eight row-pair pointers (b0 b1 b2); the b3 neighbour and the property column
  // are compile-time offsets from them
  const T* rp[8];
#pragma unroll
  for (int r = 0; r < 8; ++r) rp[r] = base + ((r >> 2) & 1) * n3 + ((r >> 1) & 1) * n2 + (r & 1) * n1;
  T w[16];
  corner_weights<T>(t0, t1, t2, t3, w);
  // full 16-byte batches, then the ninth property alone (no padding fetched).
  // From LDS a batch reads all 16 corners at once; from HBM (a path beside the
  // LDS one in the same kernels, taken rarely) it reads them in two halves, the
  // sums running on across them in the same order, so that it needs no more
  // registers than the LDS path
  constexpr int HB = FROM_LDS ? 16 : 8;
#pragma unroll
  for (int g = 0; g < 8 / VN; ++g) {
    T acc[VN];
#pragma unroll
    for (int h = 0; h < 16 / HB; ++h) {
      VT c[HB];
#pragma unroll
      for (int b = 0; b < HB; ++b) {
        const int k = HB * h + b;
        const T* q = rp[k >> 1] + (k & 1) * NP + g * VN;
        if constexpr (FROM_LDS)
          c[b] = sld<VT>(q);
        else
          c[b] = gld<VT>(q);
      }
#pragma unroll
      for (int e = 0; e < VN; ++e) {
        if (h == 0) acc[e] = w[0] * vget(c[0], e);
#pragma unroll
        for (int b = h == 0 ? 1 : 0; b < HB; ++b) acc[e] = mad<T>(w[HB * h + b], vget(c[b], e), acc[e]);
      }
      // keep the scheduler from hoisting the next loads above these sums: at
      // most one batch of rows is held in registers at a time
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int e = 0; e < VN; ++e) P[g * VN + e] = acc[e];
  }
  {
    T c[16];
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      const T* q = rp[b >> 1] + (b & 1) * NP + 8;
      if constexpr (FROM_LDS)
        c[b] = sld<T>(q);
      else
        c[b] = gld<T>(q);
    }
    T acc = w[0] * c[0];
#pragma unroll
    for (int b = 1; b < 16; ++b) acc = mad<T>(w[b], c[b], acc);
    P[8] = acc;
  }
  (void)NB;
}

// interp_core's arithmetic on one cube of the cube table (contiguous corners).
// Each 16-byte column is read in two halves of eight corners (the sums run on
// across them in the same order): half the registers of interp_core's batches,
// so the kernels that carry this path beside their LDS path do not spill
template <typename T>
PF5D_DEV void interp_cube(const T* cube, T t0, T t1, T t2, T t3, T P[9]) {
  using VT = typename Vec16<T>::type;
  constexpr int VN = Vec16<T>::n;
  T w[16];
  corner_weights<T>(t0, t1, t2, t3, w);
#pragma unroll
  for (int g = 0; g < 8 / VN; ++g) {
    T acc[VN];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      VT c[8];
#pragma unroll
      for (int b = 0; b < 8; ++b) c[b] = gld<VT>(cube + (g * 16 + 8 * h + b) * VN);
#pragma unroll
      for (int e = 0; e < VN; ++e) {
        if (h == 0) acc[e] = w[0] * vget(c[0], e);
#pragma unroll
        for (int b = h == 0 ? 1 : 0; b < 8; ++b) acc[e] = mad<T>(w[8 * h + b], vget(c[b], e), acc[e]);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int e = 0; e < VN; ++e) P[g * VN + e] = acc[e];
  }
  T c[16];
#pragma unroll
  for (int b = 0; b < 16; ++b) c[b] = gld<T>(cube + 128 + b);
  T acc = w[0] * c[0];
#pragma unroll
  for (int b = 1; b < 16; ++b) acc = mad<T>(w[b], c[b], acc);
  P[8] = acc;
}

// interpolation from the table in HBM for lower corner (i0..i3): the cube
// table when the cube is in it, the dense rows otherwise
template <typename T>
PF5D_DEV void interp_gather(const TableView<T>& tv, int i0, int i1, int i2, int i3, T t0, T t1, T t2, T t3,
                            T P[9]) {
  constexpr int NP = TablePad<T>::value;
  if (tv.cubes && i0 + i1 + i2 + i3 <= tv.cube_s) {
    interp_cube<T>(tv.cubes + (size_t)cube_rank(i0, i1, i2, i3, tv.cube_s) * CUBE_VALUES, t0, t1, t2, t3, P);
  } else {
    const int num = tv.num;
    const T* base = tv.data + ((((long)i0 * num + i1) * num + i2) * num + i3) * NP;
    interp_core<T, false>(base, num * NP, num * num * NP, num * num * num * NP, t0, t1, t2, t3, P);
  }
}

template <typename T>
PF5D_DEV void interp_linear(const TableView<T>& tv, const T X[4], T P[9]) {
  constexpr int NP = TablePad<T>::value;
  const int num = tv.num;
  int i0, i1, i2, i3;
  T t0, t1, t2, t3;
  axis_index<T>(X[0] * tv.inv_step, num, i0, t0);
  axis_index<T>(X[1] * tv.inv_step, num, i1, t1);
  axis_index<T>(X[2] * tv.inv_step, num, i2, t2);
  axis_index<T>(X[3] * tv.inv_step, num, i3, t3);
  interp_gather<T>(tv, i0, i1, i2, i3, t0, t1, t2, t3, P);
  (void)NP;
}

// A fixed box of 4^4 table nodes with lower node lo[], staged in LDS: points
// whose lower cube corner lies in its first three nodes on every axis
// interpolate from it, the others from the table in HBM (same arithmetic either
// way).  Node (q0,q1,q2,q3) starts at 16-byte slot sum_a q_a*S_a.  The strides
// are compile-time constants, so the corner offsets fold into the LDS reads, and
// they are chosen so that the 16 nodes of any 2x2x2x2 sub-box start in 16
// different slots modulo 16 (S3 odd, S2 = 2*S3, S1 = 4*S3, S0 = 8*S3 mod 16):
// when the lanes of a wavefront interpolate in neighbouring cubes, the reads of
// one corner hit distinct LDS banks instead of the same few.
constexpr int BOX4 = 4;
constexpr int box_stride(int at_least, int residue) {
  int s = at_least;
  while ((s & 15) != (residue & 15)) ++s;
  return s;
}
template <typename T>
struct BoxLayout {
  static constexpr int NCOL = TablePad<T>::value / Vec16<T>::n;  // 16-byte slots per node (odd)
  static constexpr int S3 = NCOL;
  static constexpr int S2 = box_stride(BOX4 * S3, 2 * S3);
  static constexpr int S1 = box_stride(BOX4 * S2, 4 * S3);
  static constexpr int S0 = box_stride(BOX4 * S1, 8 * S3);
  static constexpr int SLOTS = BOX4 * S0;
  static_assert(NCOL % 2 == 1, "the slot count per node must be odd");
};
template <typename T, bool WIDE>
PF5D_DEV void interp9_box4(const TableView<T>& tv, const T* box, const int lo[4], const T X[4], T P[9]) {
  constexpr int NP = TablePad<T>::value;
  const T s = ((X[0] + X[1]) + X[2]) + X[3];
  if (!(s <= tv.thresh)) {
    interp_nearest<T, WIDE>(tv, X, P);
    return;
  }
  const int num = tv.num;
  int i0, i1, i2, i3;
  T t0, t1, t2, t3;
  axis_index<T>(X[0] * tv.inv_step, num, i0, t0);
  axis_index<T>(X[1] * tv.inv_step, num, i1, t1);
  axis_index<T>(X[2] * tv.inv_step, num, i2, t2);
  axis_index<T>(X[3] * tv.inv_step, num, i3, t3);
  const int d0 = i0 - lo[0], d1 = i1 - lo[1], d2 = i2 - lo[2], d3 = i3 - lo[3];
  const bool inbox = (unsigned)d0 <= BOX4 - 2 && (unsigned)d1 <= BOX4 - 2 && (unsigned)d2 <= BOX4 - 2 &&
                     (unsigned)d3 <= BOX4 - 2;
  // A wavefront whose points all lie in the box reads it; one that mixes points
  // inside and outside (a developed field) sends every point the cube table
  // covers to the table, so it runs the 144 multiply-adds once rather than once
  // per source under opposite masks.  Box and cube table hold the same node
  // values and the sums run in the same order: the result does not depend on
  // the source.
  const bool covered = tv.cubes && i0 + i1 + i2 + i3 <= tv.cube_s;
  if (inbox && (__all(inbox) || !covered)) {
    using L = BoxLayout<T>;
    constexpr int VN = Vec16<T>::n;
    static_assert(L::S3 * VN == NP, "axis-3 stride is one padded row");
    const T* base = box + (d0 * L::S0 + d1 * L::S1 + d2 * L::S2 + d3 * L::S3) * VN;
    interp_core<T, true>(base, L::S2 * VN, L::S1 * VN, L::S0 * VN, t0, t1, t2, t3, P);
  } else {
    interp_gather<T>(tv, i0, i1, i2, i3, t0, t1, t2, t3, P);
  }
}


// Nearest table row (the reference's KD-tree query on the table rows,
// interpolation.py:203-209).  Rows are the nodes of the dense grid that hold
// values: inside the simplex and not missing from the table (build_4d_grid
// leaves NaN where a row is absent, interpolation.py:30-34).
//
// With u = X/h and f = floor(u) per axis, the boxes B_R = [f - R, f + 1 + R]^4
// are scanned, R = 0 first.  A node outside B_R is at least (R + 1) h away on
// some axis, so once the best squared distance found is at most ((R + 1) h)^2
// no node outside can be nearer.  B_0 holds the per-axis nearest node, within
// h of the point: when it is a row (the usual case) the first box settles it;
// else R = 1, 2, 4, ... until the bound holds.  Nodes are visited in
// lexicographic order and a candidate replaces the best only when strictly
// nearer, so equidistant rows resolve to the first in that order at every R.
// Missing rows are skipped by reading a candidate's first value only when it
// would become the best.
template <typename T, bool WIDE>
PF5D_DEV void nearest_scan(const TableView<T>& tv, const T X[4], const int lo[4], int a0, int a1, double& best,
                           long& bidx) {
  constexpr int NP = TablePad<T>::value;
  const int num = tv.num;
  const double h = tv.step;
#pragma unroll 1
  for (int a = a0; a < a1; ++a) {
    const int n0 = lo[0] + a;
    if (n0 < 0 || n0 >= num) continue;
#pragma unroll 1
    for (int b = a0; b < a1; ++b) {
      const int n1 = lo[1] + b;
      if (n1 < 0 || n1 >= num || n0 + n1 > num - 1) continue;
#pragma unroll 1
      for (int c = a0; c < a1; ++c) {
        const int n2 = lo[2] + c;
        if (n2 < 0 || n2 >= num || n0 + n1 + n2 > num - 1) continue;
#pragma unroll 1
        for (int e = a0; e < a1; ++e) {
          const int n3 = lo[3] + e;
          if (n3 < 0 || n3 >= num || n0 + n1 + n2 + n3 > num - 1) continue;
          const double d0 = (double)X[0] - n0 * h, d1 = (double)X[1] - n1 * h;
          const double d2 = (double)X[2] - n2 * h, d3 = (double)X[3] - n3 * h;
          const double dd = d0 * d0 + d1 * d1 + d2 * d2 + d3 * d3;
          if (dd < best) {
            const long row = (((long)n0 * num + n1) * num + n2) * num + n3;
            if (!WIDE || !__builtin_isnan((double)gld<T>(tv.data + row * NP))) {
              best = dd;
              bidx = row;
            }
          }
        }
      }
    }
  }
}

// WIDE = false, for the step kernels of engines whose tables have every row of
// the simplex: the box B_0 (16 nodes), then B_1 (4^4) when the per-axis nearest
// node lies outside the simplex, without the missing-row test -- exact for
// points within two grid steps of the simplex, which is where clamp_simplex
// keeps the step's fields (a point further out gets the nearest row of B_1, or
// NaN when B_1 holds none).  The widening search costs the potentials pass
// registers on its main path even where it never runs (7 us of 28 at 128^3
// FP32), so engines use it only when a table misses rows.
template <typename T, bool WIDE>
PF5D_DEV void interp_nearest(const TableView<T>& tv, const T X[4], T P[9]) {
  constexpr int NP = TablePad<T>::value;
  const int num = tv.num;
  const double h = tv.step;
  int lo[4];
  bool finite = true, near_in = true;
  int near_sum = 0;
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    const double u = (double)X[d] / h;
    if constexpr (WIDE) finite = finite && __builtin_isfinite(u);  // (NaN distances find no row anyway)
    // (this form of the clamp keeps the potentials pass's registers at their
    // best allocation; a NaN or huge u converts to 0 or the largest int on gfx950)
    int f = (u <= 0.0) ? 0 : (int)u;
    if (f > num - 1) f = num - 1;
    if constexpr (WIDE) f = (u == u) ? f : 0;
    lo[d] = f - 1;
    // per-axis nearest of the candidates f, f + 1 (ties to f, as the scan keeps the first)
    const double df = (double)X[d] - f * h, dg = (double)X[d] - (f + 1) * h;
    const int r = (f + 1 < num && dg * dg < df * df) ? f + 1 : f;
    near_sum += r;
  }
  near_in = near_sum <= num - 1;
  // (infinity, not a large finite bound: its low word is zero, so it is an
  // instruction literal instead of a register pair held across the caller's loop)
  double best = __builtin_huge_val();
  long bidx = -1;
  if (finite) {
    nearest_scan<T, WIDE>(tv, X, lo, 1, 3, best, bidx);
    if constexpr (WIDE) {
      int R = 0;
#pragma unroll 1
      while (!(bidx >= 0 && best <= (double)((R + 1) * (R + 1)) * (h * h)) && R <= 2 * num) {
        R = (R == 0) ? 1 : 2 * R;
        best = __builtin_huge_val();
        bidx = -1;
        nearest_scan<T, WIDE>(tv, X, lo, 1 - R, 3 + R, best, bidx);
      }
    } else {
      if (!near_in) {
        best = __builtin_huge_val();
        bidx = -1;
        nearest_scan<T, WIDE>(tv, X, lo, 0, 4, best, bidx);
      }
    }
  }
  if (bidx < 0) {
#pragma unroll
    for (int p = 0; p < 9; ++p) P[p] = T(__builtin_nan(""));
    return;
  }
  T g[NP];
#pragma unroll
  for (int p = 0; p < 9; ++p) g[p] = gld<T>(tv.data + bidx * NP + p);
#pragma unroll
  for (int p = 0; p < 9; ++p) P[p] = g[p];
}

template <typename T>
PF5D_DEV void interp9(const TableView<T>& tv, const T X[4], T P[9]) {
  const T s = ((X[0] + X[1]) + X[2]) + X[3];
  if (s <= tv.thresh)
    interp_linear<T>(tv, X, P);
  else
    interp_nearest<T, true>(tv, X, P);
}

}  // namespace pf5d
