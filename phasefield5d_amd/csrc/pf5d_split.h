// The stencil stage K1: interpolation + kappa lap(X) + elastic psi -> mu, Mob
// -> face fluxes -> dX/dt and max |dX/dt| per simulation, in one of two forms
// that give the same bits.
//
// Two-pass form:
//   KP  k_potentials : per cell: interpolation + kappa lap(X) + elastic psi
//                      -> mu (4 values) and Mob (5 values) in HBM; the table
//                      nodes around the tile's centre cell (4^4) are staged in
//                      LDS and serve every point whose cube lies among them
//   KF  k_fluxmarch  : marching along axis 0 over tiles of rows: the face
//                      fluxes from (X, mu, Mob), each face once -> dX/dt
//                      (k_fluxdiv: the same per plane, each face from both
//                      cells, for the per-plane tilings)
// Fused marching form (k_stencilmarch): the marching pass computes mu and Mob
// of its tile's inner rows itself, one plane ahead, into LDS, and reads them
// from HBM only for the rows and planes on its tiles' boundaries, which KP
// computes.
//
// Work items of KP are (sim, x planes xa..xb-1, y0, z0, ty, tz) tiles of at
// most NT cells per plane: thread t handles cells (x, y0 + t / tz, z0 + t % tz)
// for x in [xa, xb), all served by one staged node box.  Items are listed with
// x fastest and every block takes a contiguous range of them, so a block walks
// consecutive planes of one tile and finds its x neighbours in cache.  A face
// flux is evaluated with the lower cell first, so a face evaluated once (the
// marching passes) or by both its cells (k_fluxdiv) gives the same value, and
// the divergence telescopes exactly.
#pragma once
#include "pf5d_common.h"
#include "pf5d_flux.h"
#include "pf5d_fused.h"
#include "pf5d_interp.h"

namespace pf5d {

// mu of one cell from its interpolated properties P, the sum acc of its
// neighbours' X (x, then y, then z: each up + down, as calculate_laplacian
// adds them) and the elastic potential psi: (acc - 2 dim x) / h^2, then
// P - L kappa (+ psi l) (simulate.py:253-255)
template <typename T>
PF5D_DEV V4<T> cell_mu(const SimDev<T>& sd, const V4<T>& xc, const T acc[4], const T P[9], T psi) {
  const T cdim = T(2 * sd.dim);
  V4<T> mu;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const T L = mad<T>(-cdim, xc.v[a], acc[a]) * sd.inv_dx2;
    T m = mad<T>(-L, sd.kappa[a], P[a]);
    if (sd.elastic) m = mad<T>(psi, sd.l[a], m);
    mu.v[a] = m;
  }
  return mu;
}

// Stages the 4^4 node box of an item in LDS (BOX, BoxLayout<T>::SLOTS
// 16-byte slots): the box around the centre cell of the item's planes (the
// anchor), lower nodes one below its lower cube corners, so corners within +-1
// of the anchor's are covered; lo[] receives the box's lower node.  Every
// thread of the block calls it; the caller synchronises before reading BOX.
template <typename T, int NT>
PF5D_DEV void stage_box4(const SimDev<T>& sd, const K1Item& item, T* BOX, int lo[4], int tid) {
  constexpr int NP = TablePad<T>::value;
  constexpr int NBOX = BOX4 * BOX4 * BOX4 * BOX4;
  using VT = typename Vec16<T>::type;
  using L = BoxLayout<T>;
  constexpr int NCOL = L::NCOL;  // 16-byte columns per table row
  const int num = sd.tab.num;
  {
    const long ca = ((long)(item.xa + (item.xb - 1 - item.xa) / 2) * sd.n1 + item.y0 + item.ty / 2) * sd.n2 +
                    item.z0 + item.tz / 2;
    const V4<T> xa = ld4g<T>(sd.X + 4 * ca);
    // (the box position only decides which points read LDS: clamped in
    // floating point first, so NaN lands on 0 and the conversion is in range;
    // every thread computes the same box, and readfirstlane keeps it in scalar
    // registers, where the per-point box tests read it)
    const float top = (float)(num - BOX4);
#pragma unroll
    for (int a = 0; a < 4; ++a)
      lo[a] = __builtin_amdgcn_readfirstlane((int)fminf(fmaxf((float)(xa.v[a] * sd.tab.inv_step) - 1.0f, 0.0f), top));
  }
  // one node per thread, all of its columns: the box's first row (64-bit),
  // then offsets within the box (below 4 num^3 NP, 32-bit)
  const T* row0 = sd.tab.data + ((((long)lo[0] * num + lo[1]) * num + lo[2]) * num + lo[3]) * NP;
  const int m2 = num * num, m3 = m2 * num;
#pragma unroll 1
  for (int node = tid; node < NBOX; node += NT) {
    const int q0 = node >> 6, q1 = (node >> 4) & 3, q2 = (node >> 2) & 3, q3 = node & 3;
    const T* src = row0 + (q0 * m3 + q1 * m2 + q2 * num + q3) * NP;
    VT* dst = reinterpret_cast<VT*>(BOX) + (q0 * L::S0 + q1 * L::S1 + q2 * L::S2 + q3 * L::S3);
    VT v[NCOL];
#pragma unroll
    for (int col = 0; col < NCOL; ++col) v[col] = gld<VT>(src + col * Vec16<T>::n);
#pragma unroll
    for (int col = 0; col < NCOL; ++col) dst[col] = v[col];
  }
}

// cell index of the tile thread, and whether it has a cell
template <typename T>
PF5D_DEV bool tile_cell(const SimDev<T>& sd, const K1Item& item, int tid, float rtz, int& gx, int& gy, int& gz) {
  int j, k;
  fdivmod(tid, item.tz, rtz, j, k);
  gx = item.xa;
  gy = item.y0 + j;
  gz = item.z0 + k;
  return j < item.ty;
}

// MINW: minimum waves per SIMD the register allocation must allow; WIDE: the
// nearest-row search for tables with missing rows (interp_nearest)
template <typename T, int NT, int MINW, bool WIDE>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(MINW, 8)))
k_potentials(const SimDev<T>* __restrict__ sims, const SimState* __restrict__ st, const K1Item* __restrict__ items,
             int nitems) {
  using L = BoxLayout<T>;
  __shared__ __align__(16) T BOX[L::SLOTS * Vec16<T>::n];
  int ib, ie;
  block_range(nitems, ib, ie);
  // the thread index is rebuilt per item from the wavefront index (a scalar)
  // and the lane count below it (every lane is active at the loop's top), so no
  // vector register holds it across the items
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
  for (int it = ib; it < ie; ++it) {
    const K1Item item = items[it];
    if (!(st[item.sim].status & ST_ACTIVE)) continue;  // uniform per block
    // an empty asm keeps address arithmetic from being hoisted out of the item
    // loop (holding it across items costs more registers than recomputing it)
    unsigned ones = ~0u;
    asm volatile("" : "+s"(ones));  // (opaque, so the lane count is not hoisted out of the loop)
    int tid = wave * 64 + (int)__builtin_amdgcn_mbcnt_hi(ones, __builtin_amdgcn_mbcnt_lo(ones, 0u));
    asm volatile("" : "+v"(tid));
    const SimDev<T>& sd = sims[item.sim];
    const int n0 = sd.n0, n1 = sd.n1, n2 = sd.n2;
    const bool hx = sd.dim == 3, hy = sd.dim >= 2;
    int gx, gy, gz;
    const bool valid = tile_cell<T>(sd, item, tid, frcp(item.tz), gx, gy, gz);
    // byte offsets of the cell in the (ncell, 4) and (ncell) fields, 32-bit
    // (the engine admits at most 2^32 / (4 sizeof(T)) cells; the unsigned sums
    // below wrap modulo 2^32 to the in-range offsets they stand for)
    constexpr unsigned B4 = sizeof(V4<T>), B1 = sizeof(T);
    unsigned c = ((unsigned)gx * n1 + gy) * n2 + gz;
    V4<T> xc{};
    if (valid) xc = gldo<V4<T>>(sd.X, c * B4);
    int lo[4];
    stage_box4<T, NT>(sd, item, BOX, lo, tid);
    __syncthreads();
    // the item's planes xa .. xb-1 share the staged box: one cell per plane.
    // FP32 walks up in x with the cell below (xm) and the cell itself carried
    // over from the previous plane; FP64 reloads them (carrying them costs it
    // registers the fallback paths need)
    constexpr bool carry = sizeof(T) == 4;
    const unsigned sx = (unsigned)n1 * n2;
    // neighbour offsets in cells (periodic): x below and above, y, z
    const unsigned dxd = gx > 0 ? 0u - sx : (unsigned)(n0 - 1) * sx;
    const int dyu = (wrap(gy + 1, n1) - gy) * n2, dyd = (wrap(gy - 1, n1) - gy) * n2;
    const int dzu = wrap(gz + 1, n2) - gz, dzd = wrap(gz - 1, n2) - gz;
    V4<T> xm{};
    if (carry && valid && hx) xm = gldo<V4<T>>(sd.X, (c + dxd) * B4);
#pragma unroll 1
    for (gx = item.xa; valid && gx < item.xb; ++gx, c += sx) {
      const unsigned c4 = c * B4;
      const unsigned dxu = gx + 1 < n0 ? sx : sx - (unsigned)n0 * sx;
      if (!carry && gx > item.xa) xc = gldo<V4<T>>(sd.X, c4);
      V4<T> xp{};
      if (carry && hx) xp = gldo<V4<T>>(sd.X, (c + dxu) * B4);
      T P[9];
      interp9_box4<T, WIDE>(sd.tab, BOX, lo, xc.v, P);
      // neighbour sums in the reference order: x, then y, then z (each up + down)
      T acc[4];
      bool first = true;
      if (hx) {
        if constexpr (carry) {
#pragma unroll
          for (int a = 0; a < 4; ++a) acc[a] = xp.v[a] + xm.v[a];
        } else {
          const V4<T> u = gldo<V4<T>>(sd.X, (c + dxu) * B4);
          const V4<T> d = gldo<V4<T>>(sd.X, (c + (gx > 0 ? 0u - sx : (unsigned)(n0 - 1) * sx)) * B4);
#pragma unroll
          for (int a = 0; a < 4; ++a) acc[a] = u.v[a] + d.v[a];
        }
        first = false;
      }
      if (hy) {
        const V4<T> u = gldo<V4<T>>(sd.X, c4 + dyu * (int)B4), d = gldo<V4<T>>(sd.X, c4 + dyd * (int)B4);
#pragma unroll
        for (int a = 0; a < 4; ++a) acc[a] = first ? (u.v[a] + d.v[a]) : ((acc[a] + u.v[a]) + d.v[a]);
        first = false;
      }
      {
        const V4<T> u = gldo<V4<T>>(sd.X, c4 + dzu * (int)B4);
        const V4<T> d = gldo<V4<T>>(sd.X, c4 + dzd * (int)B4);
#pragma unroll
        for (int a = 0; a < 4; ++a) acc[a] = first ? (u.v[a] + d.v[a]) : ((acc[a] + u.v[a]) + d.v[a]);
      }
      const T psi = sd.elastic ? gldo<T>(sd.psi, c * B1) : T(0);
      const V4<T> mu = cell_mu<T>(sd, xc, acc, P, psi);
      V4<T> mb;
#pragma unroll
      for (int a = 0; a < 4; ++a) mb.v[a] = P[4 + a];
      gst_field_o<T, V4<T>>(sd.mu, c4, mu);
      gst_field_o<T, V4<T>>(sd.mb, c4, mb);
      gst_field_o<T, T>(sd.mb1, c * B1, P[8]);
      if (carry) {
        xm = xc;
        xc = xp;
      }
    }
    __syncthreads();  // BOX is reused by the next item
  }
}

// X, mu and Mob of one cell
template <typename T>
struct CellV {
  V4<T> X, mu;
  T M[5];
};

template <typename T>
PF5D_DEV CellV<T> load_cell(const SimDev<T>& sd, long c) {
  CellV<T> r;
  r.X = ld4g<T>(sd.X + 4 * c);
  r.mu = ld4g<T>(sd.mu + 4 * c);
  const V4<T> m = ld4g<T>(sd.mb + 4 * c);
#pragma unroll
  for (int a = 0; a < 4; ++a) r.M[a] = m.v[a];
  r.M[4] = gld<T>(sd.mb1 + c);
  return r;
}
// the same for a cell index below 2^32 / (4 sizeof(T)), as byte offsets from
// the fields' bases (see gldo)
template <typename T>
PF5D_DEV CellV<T> load_cell32(const SimDev<T>& sd, unsigned c) {
  CellV<T> r;
  const unsigned o4 = c * (unsigned)sizeof(V4<T>), o1 = c * (unsigned)sizeof(T);
  r.X = gldo<V4<T>>(sd.X, o4);
  r.mu = gldo<V4<T>>(sd.mu, o4);
  const V4<T> m = gldo<V4<T>>(sd.mb, o4);
#pragma unroll
  for (int a = 0; a < 4; ++a) r.M[a] = m.v[a];
  r.M[4] = gldo<T>(sd.mb1, o1);
  return r;
}

// flux through the face from cell a (lower) to cell b (upper)
template <typename T>
PF5D_DEV void face_ab(const CellV<T>& a, const CellV<T>& b, T inv_dx, T out[4]) {
  step_face_flux<T>(a.X.v, b.X.v, a.M, b.M, a.mu.v, b.mu.v, inv_dx, out);
}

// the X, mu and Mob of lane (lane - 1) (dir = -1) or (lane + 1) (dir = +1) of the wavefront
template <typename T>
PF5D_DEV CellV<T> shfl_cell(const CellV<T>& v, int dir) {
  CellV<T> r;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    r.X.v[a] = dir < 0 ? __shfl_up(v.X.v[a], 1, 64) : __shfl_down(v.X.v[a], 1, 64);
    r.mu.v[a] = dir < 0 ? __shfl_up(v.mu.v[a], 1, 64) : __shfl_down(v.mu.v[a], 1, 64);
  }
#pragma unroll
  for (int m = 0; m < 5; ++m) r.M[m] = dir < 0 ? __shfl_up(v.M[m], 1, 64) : __shfl_down(v.M[m], 1, 64);
  return r;
}

// LOADS: 0 = the two neighbours of an axis loaded when the axis is evaluated;
// 1 = all seven cells requested before any face is evaluated (one memory round
// trip, more registers); 2 = as 1, but the z neighbours come from the adjacent
// lanes of the wavefront (consecutive z), loaded only at row and wavefront ends
template <typename T, int NT, int LOADS>
__global__ void __launch_bounds__(NT) k_fluxdiv(const SimDev<T>* __restrict__ sims, const SimState* __restrict__ st,
                                                T* __restrict__ xmax_acc, const K1Item* __restrict__ items,
                                                int nitems) {
  __shared__ T RED[NT / 64];
  const int tid = threadIdx.x;
  // a simulation's items are adjacent, so the block max is flushed with one
  // atomic per simulation the block's range touches
  int ib, ie;
  block_range(nitems, ib, ie);
  T tmax = T(0);
  int cur = -1;
  auto flush = [&](int sim) {
    const T m = wave_max<T>(tmax);
    if ((tid & 63) == 0) RED[tid >> 6] = m;
    __syncthreads();
    if (tid == 0) {
      T r = RED[0];
      for (int w = 1; w < NT / 64; ++w) r = (RED[w] > r) ? RED[w] : r;
      atomic_max_nonneg(&xmax_acc[sim], r);
    }
    __syncthreads();
    tmax = T(0);
  };
  for (int it = ib; it < ie; ++it) {
    const K1Item item = items[it];
    if (item.sim != cur) {
      if (cur >= 0) flush(cur);
      cur = item.sim;
    }
    if (!(st[item.sim].status & ST_ACTIVE)) continue;  // uniform per block
    const SimDev<T>& sd = sims[item.sim];
    const int n0 = sd.n0, n1 = sd.n1, n2 = sd.n2;
    const bool hx = sd.dim == 3, hy = sd.dim >= 2;
    int gx, gy, gz;
    const bool valid = tile_cell<T>(sd, item, tid, frcp(item.tz), gx, gy, gz);
    if constexpr (LOADS < 2) {
      if (!valid) continue;
    } else if (!valid) {
      // lanes past the tile's rows take part in the lane exchange on the tile's first cell
      gy = item.y0;
      gz = item.z0;
    }
    const long c = ((long)gx * n1 + gy) * n2 + gz;
    const T inv_dx = sd.inv_dx, dscale = step_div_scale<T>(inv_dx);
    const long sx = (long)n1 * n2;
    const long ixu = hx ? c + (gx + 1 < n0 ? sx : sx - (long)n0 * sx) : c;
    const long ixd = hx ? c + (gx > 0 ? -sx : (long)(n0 - 1) * sx) : c;
    const long iyu = hy ? c + (long)(wrap(gy + 1, n1) - gy) * n2 : c;
    const long iyd = hy ? c + (long)(wrap(gy - 1, n1) - gy) * n2 : c;
    const long izu = c + (wrap(gz + 1, n2) - gz), izd = c + (wrap(gz - 1, n2) - gz);
    const CellV<T> me = load_cell<T>(sd, c);
    // absent axes read the cell itself (not used)
    CellV<T> pre[6];
    if constexpr (LOADS >= 1) {
      pre[0] = load_cell<T>(sd, ixu);
      pre[1] = load_cell<T>(sd, ixd);
      pre[2] = load_cell<T>(sd, iyu);
      pre[3] = load_cell<T>(sd, iyd);
    }
    if constexpr (LOADS == 1) {
      pre[4] = load_cell<T>(sd, izu);
      pre[5] = load_cell<T>(sd, izd);
    } else if constexpr (LOADS == 2) {
      const int lane = tid & 63, k = gz - item.z0;
      pre[4] = shfl_cell<T>(me, +1);
      pre[5] = shfl_cell<T>(me, -1);
      if (!(k < item.tz - 1 && lane < 63)) pre[4] = load_cell<T>(sd, izu);
      if (!(k > 0 && lane > 0)) pre[5] = load_cell<T>(sd, izd);
    }
    T acc[4];
    bool first = true;
    // divergence in the reference order: x, y, z, each (F+ - F-) added in turn
    auto axis = [&](int a0, long iu, long id) {
      T fp[4], fm[4];
      if constexpr (LOADS >= 1) {
        face_ab<T>(me, pre[a0], inv_dx, fp);
        face_ab<T>(pre[a0 + 1], me, inv_dx, fm);
      } else {
        const CellV<T> up = load_cell<T>(sd, iu), dn = load_cell<T>(sd, id);
        face_ab<T>(me, up, inv_dx, fp);
        face_ab<T>(dn, me, inv_dx, fm);
      }
#pragma unroll
      for (int a = 0; a < 4; ++a) acc[a] = first ? (fp[a] - fm[a]) : ((acc[a] + fp[a]) - fm[a]);
      first = false;
    };
    if (hx) axis(0, ixu, ixd);
    if (hy) axis(2, iyu, iyd);
    axis(4, izu, izd);
    V4<T> o;
#pragma unroll
    for (int a = 0; a < 4; ++a) o.v[a] = acc[a] * dscale;
    if (valid) {
#pragma unroll
      for (int a = 0; a < 4; ++a) tmax = max_ignore_nan<T>(tmax, fabs(o.v[a]));
      st4g<T>(sd.dX + 4 * c, o);
    }
  }
  if (cur >= 0) flush(cur);
}

// 16 bytes of T, the unit of the LDS exchange below
template <typename T>
struct alignas(16) Q16 {
  T v[16 / sizeof(T)];
};

// cell values of a lane, packed for the LDS exchange: X, mu, Mob of each cell
template <typename T, int CPL, int QC>
PF5D_DEV void put_cells(Q16<T> (*dst)[64], int lane, const CellV<T> (&c)[CPL]) {
  constexpr int QN = 16 / sizeof(T);
  T p[QC * QN];
#pragma unroll
  for (int j = 0; j < CPL; ++j) {
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      p[13 * j + a] = c[j].X.v[a];
      p[13 * j + 4 + a] = c[j].mu.v[a];
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) p[13 * j + 8 + m] = c[j].M[m];
  }
#pragma unroll
  for (int i = 13 * CPL; i < QC * QN; ++i) p[i] = T(0);
#pragma unroll
  for (int q = 0; q < QC; ++q) {
    Q16<T> v;
#pragma unroll
    for (int e = 0; e < QN; ++e) v.v[e] = p[q * QN + e];
    sst<Q16<T>>(&dst[q][lane], v);
  }
}

template <typename T, int CPL, int QC>
PF5D_DEV void get_cells(const Q16<T> (*src)[64], int lane, CellV<T> (&c)[CPL]) {
  constexpr int QN = 16 / sizeof(T);
  T p[QC * QN];
#pragma unroll
  for (int q = 0; q < QC; ++q) {
    const Q16<T> v = sld<Q16<T>>(&src[q][lane]);
#pragma unroll
    for (int e = 0; e < QN; ++e) p[q * QN + e] = v.v[e];
  }
#pragma unroll
  for (int j = 0; j < CPL; ++j) {
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      c[j].X.v[a] = p[13 * j + a];
      c[j].mu.v[a] = p[13 * j + 4 + a];
    }
#pragma unroll
    for (int m = 0; m < 5; ++m) c[j].M[m] = p[13 * j + 8 + m];
  }
}

// Whole-wavefront rotations by one lane (DPP, a VALU move: no LDS round trip):
// rot_next gives lane i the value of lane (i + 1) mod 64, rot_prev that of lane
// (i - 1) mod 64.  Every lane must be active.
template <int CTRL>
PF5D_DEV float dpp_rot(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
PF5D_DEV double dpp_rot(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
constexpr int DPP_WAVE_ROL1 = 0x134;  // wave_rol:1
constexpr int DPP_WAVE_ROR1 = 0x13C;  // wave_ror:1
template <typename T>
PF5D_DEV T rot_next(T v) { return dpp_rot<DPP_WAVE_ROL1>(v); }
template <typename T>
PF5D_DEV T rot_prev(T v) { return dpp_rot<DPP_WAVE_ROR1>(v); }
template <typename T>
PF5D_DEV CellV<T> rot_next_cell(const CellV<T>& v) {
  CellV<T> r;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    r.X.v[a] = rot_next<T>(v.X.v[a]);
    r.mu.v[a] = rot_next<T>(v.mu.v[a]);
  }
#pragma unroll
  for (int m = 0; m < 5; ++m) r.M[m] = rot_next<T>(v.M[m]);
  return r;
}

// X, mu, Mob of lane `src` of the wavefront
template <typename T>
PF5D_DEV CellV<T> shfl_cell_from(const CellV<T>& v, int src) {
  CellV<T> r;
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    r.X.v[a] = __shfl(v.X.v[a], src, 64);
    r.mu.v[a] = __shfl(v.mu.v[a], src, 64);
  }
#pragma unroll
  for (int m = 0; m < 5; ++m) r.M[m] = __shfl(v.M[m], src, 64);
  return r;
}

// Flux pass marching along axis 0: every face is evaluated once.
//
// An item is a tile of rows (one per wavefront, at most NT/64) of up to 64*CPL
// cells on axis 2, and a range [xa, xb) of planes on axis 0; lane i holds the
// cells CPL*i .. CPL*i+CPL-1 of its wavefront's row.  Per plane:
// - the x face between the plane and the previous one in marching order (up
//   or down the range, by build) was evaluated there and kept in registers, so
//   one x face is evaluated per cell, plus one per cell for the first plane;
// - the y face above a row uses the next row's cells, taken from LDS (the last
//   row loads the row after the tile); the y face below a row is the face above
//   the previous row, taken from LDS, and only the first row evaluates its own
//   (or, when the tile spans the periodic axis, takes the last row's);
// - the z face above a cell uses the next cell, taken from the same lane or by
//   a lane shuffle; the z face below is the face above the previous cell, and
//   only the row's first cell evaluates its own when the tile does not span the
//   periodic row.
// Faces are computed with the lower cell first, so a face evaluated once gives
// the same bits as the reference's two evaluations (the face flux is symmetric),
// and the divergence is summed in the reference order.
//
// With ndt > 0 the block that finishes last also runs the dt controller of
// simulations dt_list[0 .. ndt-1] (0 .. ndt-1 without a list; the step's DT
// stage, without a launch of its own):
// every block counts itself in *done after its maxima are folded in; the last
// one to count resets the counter for the next launch.
// MINW: minimum waves per SIMD the register allocation must allow
template <typename T, int NT, int CPL, int MINW>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(MINW, 8))) k_fluxmarch(const SimDev<T>* __restrict__ sims, SimState* __restrict__ st,
                                                  T* __restrict__ xmax_acc, const K1Item* __restrict__ items,
                                                  int nitems, const double* __restrict__ part,
                                                  const int* __restrict__ part_range, int ndt,
                                                  const int* __restrict__ dt_list, unsigned* __restrict__ done) {
  constexpr int NW = NT / 64;
  constexpr int QN = 16 / sizeof(T);
  constexpr int QC = (13 * CPL + QN - 1) / QN;  // 16-byte words per lane: cells
  constexpr int QF = (4 * CPL + QN - 1) / QN;   // 16-byte words per lane: one face per cell
  __shared__ Q16<T> SC[NW][QC][64];
  __shared__ Q16<T> SF[NW][QF][64];
  __shared__ T RED[NW];
  const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  int ib, ie;
  block_range(nitems, ib, ie);
  T tmax = T(0);
  int cur = -1;
  unsigned long long folded = 0;  // previous bits returned by the maxima's atomics
  auto flush = [&](int sim) {
    const T m = wave_max<T>(tmax);
    if (lane == 0) RED[w] = m;
    __syncthreads();
    if (tid == 0) {
      T r = RED[0];
      for (int v = 1; v < NW; ++v) r = (RED[v] > r) ? RED[v] : r;
      folded |= atomic_max_nonneg(&xmax_acc[sim], r);
    }
    __syncthreads();
    tmax = T(0);
  };
  for (int it = ib; it < ie; ++it) {
    const K1Item item = items[it];
    if (item.sim != cur) {
      if (cur >= 0) flush(cur);
      cur = item.sim;
    }
    if (!(st[item.sim].status & ST_ACTIVE)) continue;  // uniform per block
    const SimDev<T>& sd = sims[item.sim];
    const int n0 = sd.n0, n1 = sd.n1, n2 = sd.n2;
    const bool hx = sd.dim == 3, hy = sd.dim >= 2;
    const T inv_dx = sd.inv_dx, dscale = step_div_scale<T>(inv_dx);
    // wavefronts past the tile's rows work on its first row and store nothing
    const bool row_on = w < item.ty;
    const int gy = item.y0 + (row_on ? w : 0);
    const int tz = item.tz;
    const bool zwrap = tz == n2, ywrap = item.ty == n1;
    const int lend = (tz - 1) / CPL, jend = (tz - 1) - lend * CPL;  // lane and slot of the row's last cell
    unsigned kz[CPL];
    bool kon[CPL];
#pragma unroll
    for (int j = 0; j < CPL; ++j) {
      const int k = CPL * lane + j;
      kon[j] = row_on && k < tz;
      kz[j] = item.z0 + min(k, tz - 1);
    }
    // cell indices in 32 bits (the engine admits at most 2^32 / (4 sizeof(T)) cells)
    auto row_base = [&](int x, int y) { return ((unsigned)x * n1 + y) * n2; };
    auto load_lane = [&](int x, int y, CellV<T>(&c)[CPL]) {
      const unsigned r = row_base(x, y);
#pragma unroll
      for (int j = 0; j < CPL; ++j) c[j] = load_cell32<T>(sd, r + kz[j]);
    };
    // FP32 two-wave builds march down from the run's last plane: the planes the
    // potentials pass wrote last are the first this pass reads (1 us at 128^3).
    // Either way the x face carried from plane to plane (fxc) is the one between
    // the plane and the previous one in marching order, evaluated lower cell
    // first, so both directions give the same bits.  (FP64 and the three-wave
    // builds march up: down, they spill.)
    constexpr bool down = sizeof(T) == 4 && MINW < 3;
    CellV<T> me[CPL];
    T fxc[CPL][4];
    load_lane(down ? item.xb - 1 : item.xa, gy, me);
    if (hx) {
      CellV<T> pv[CPL];
      if constexpr (down) {
        load_lane(item.xb < n0 ? item.xb : 0, gy, pv);
#pragma unroll
        for (int j = 0; j < CPL; ++j) face_ab<T>(me[j], pv[j], inv_dx, fxc[j]);
      } else {
        load_lane(item.xa > 0 ? item.xa - 1 : n0 - 1, gy, pv);
#pragma unroll
        for (int j = 0; j < CPL; ++j) face_ab<T>(pv[j], me[j], inv_dx, fxc[j]);
      }
    }
    // FP32: the tile's edge wavefronts load the halo row beyond it at the start
    // of each plane, so the load overlaps the exchange instead of stalling the
    // block after its first barrier (a tile of one row loads both in place).
    // FP64, and FP32 built for three waves per SIMD, load them in place: the
    // extra row of registers made those builds spill
    constexpr bool prefetch_halo = sizeof(T) == 4 && MINW < 3;
    const bool pre_up = prefetch_halo && hy && !ywrap && item.ty > 1 && w == item.ty - 1;
    const bool pre_dn = prefetch_halo && hy && !ywrap && item.ty > 1 && w == 0;
    for (int x = down ? item.xb - 1 : item.xa; down ? x >= item.xa : x < item.xb; x += down ? -1 : 1) {
      CellV<T> hal[CPL];
      if (pre_up)
        load_lane(x, gy + 1 < n1 ? gy + 1 : 0, hal);
      else if (pre_dn)
        load_lane(x, gy > 0 ? gy - 1 : n1 - 1, hal);
      CellV<T> nx[CPL];  // the next plane in marching order
      if (hx) {
        if constexpr (down)
          load_lane(x > 0 ? x - 1 : n0 - 1, gy, nx);
        else
          load_lane(x + 1 < n0 ? x + 1 : 0, gy, nx);
      }
      T fyu[CPL][4], fyd[CPL][4];
      if (hy) {
        put_cells<T, CPL, QC>(SC[w], lane, me);
        __syncthreads();
        {
          CellV<T> up[CPL];
          if (w + 1 < item.ty)
            get_cells<T, CPL, QC>(SC[w + 1], lane, up);
          else if (ywrap)
            get_cells<T, CPL, QC>(SC[0], lane, up);
          else if (pre_up) {
#pragma unroll
            for (int j = 0; j < CPL; ++j) up[j] = hal[j];
          } else
            load_lane(x, gy + 1 < n1 ? gy + 1 : 0, up);
#pragma unroll
          for (int j = 0; j < CPL; ++j) face_ab<T>(me[j], up[j], inv_dx, fyu[j]);
        }
        {
          T p[QF * QN];
#pragma unroll
          for (int j = 0; j < CPL; ++j)
#pragma unroll
            for (int a = 0; a < 4; ++a) p[4 * j + a] = fyu[j][a];
#pragma unroll
          for (int i = 4 * CPL; i < QF * QN; ++i) p[i] = T(0);
#pragma unroll
          for (int q = 0; q < QF; ++q) {
            Q16<T> v;
#pragma unroll
            for (int e = 0; e < QN; ++e) v.v[e] = p[q * QN + e];
            sst<Q16<T>>(&SF[w][q][lane], v);
          }
        }
        __syncthreads();
        if (w > 0 || ywrap) {
          const int src = (w > 0) ? w - 1 : item.ty - 1;
          T p[QF * QN];
#pragma unroll
          for (int q = 0; q < QF; ++q) {
            const Q16<T> v = sld<Q16<T>>(&SF[src][q][lane]);
#pragma unroll
            for (int e = 0; e < QN; ++e) p[q * QN + e] = v.v[e];
          }
#pragma unroll
          for (int j = 0; j < CPL; ++j)
#pragma unroll
            for (int a = 0; a < 4; ++a) fyd[j][a] = p[4 * j + a];
        } else {
          CellV<T> dn[CPL];
          if (pre_dn) {
#pragma unroll
            for (int j = 0; j < CPL; ++j) dn[j] = hal[j];
          } else {
            load_lane(x, gy > 0 ? gy - 1 : n1 - 1, dn);
          }
#pragma unroll
          for (int j = 0; j < CPL; ++j) face_ab<T>(dn[j], me[j], inv_dx, fyd[j]);
        }
      }
      // z: the face above each cell, then the face below from the cell before
      T fzu[CPL][4], fzd[CPL][4];
      if (tz == 64 * CPL && zwrap) {
        // a periodic row of 64 * CPL cells fills the wavefront (uniform over the
        // block): the cell after the lane's last is the next lane's first, a
        // rotation; the same faces as the general form below
        const CellV<T> nb = rot_next_cell<T>(me[0]);
#pragma unroll
        for (int j = 0; j + 1 < CPL; ++j) face_ab<T>(me[j], me[j + 1], inv_dx, fzu[j]);
        face_ab<T>(me[CPL - 1], nb, inv_dx, fzu[CPL - 1]);
#pragma unroll
        for (int j = 1; j < CPL; ++j)
#pragma unroll
          for (int a = 0; a < 4; ++a) fzd[j][a] = fzu[j - 1][a];
#pragma unroll
        for (int a = 0; a < 4; ++a) fzd[0][a] = rot_prev<T>(fzu[CPL - 1][a]);
      } else {
        // the cell after the lane's last one: the next lane's first, or, for the
        // row's last cell, the row's first (periodic tile) or the cell after the tile
        const bool ends = lane == lend;
        // a row of 64 * CPL cells fills the wavefront: the next lane is a rotation
        // (uniform over the block: the tile's row length)
        const bool full = tz == 64 * CPL;
        CellV<T> nb = full ? rot_next_cell<T>(me[0]) : shfl_cell_from<T>(me[0], (zwrap && ends) ? 0 : ((lane + 1) & 63));
        if (!zwrap && ends) nb = load_cell32<T>(sd, row_base(x, gy) + (item.z0 + tz < n2 ? item.z0 + tz : 0));
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
          const bool last = ends && j == jend;
          const CellV<T>& nxt = (j + 1 < CPL) ? me[j + 1 < CPL ? j + 1 : 0] : nb;
          CellV<T> u;
          u.X = last ? nb.X : nxt.X;
          u.mu = last ? nb.mu : nxt.mu;
#pragma unroll
          for (int m = 0; m < 5; ++m) u.M[m] = last ? nb.M[m] : nxt.M[m];
          face_ab<T>(me[j], u, inv_dx, fzu[j]);
        }
#pragma unroll
        for (int j = 1; j < CPL; ++j)
#pragma unroll
          for (int a = 0; a < 4; ++a) fzd[j][a] = fzu[j - 1][a];
        // below the lane's first cell: the previous lane's last face; lane 0 of a
        // periodic tile: the row's last face (the lane holding the row's last
        // cell offers that one, its later slots being outside the row)
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          T v = fzu[CPL - 1][a];
#pragma unroll
          for (int j = 0; j + 1 < CPL; ++j) v = (ends && jend == j) ? fzu[j][a] : v;
          fzd[0][a] = full ? rot_prev<T>(v) : __shfl(v, (zwrap && lane == 0) ? lend : ((lane + 63) & 63), 64);
        }
        if (lane == 0 && !zwrap) {
          const CellV<T> dn = load_cell32<T>(sd, row_base(x, gy) + (item.z0 > 0 ? item.z0 - 1 : n2 - 1));
          face_ab<T>(dn, me[0], inv_dx, fzd[0]);
        }
      }
      // the x face with the next plane (carried on to it); with the carried one,
      // the faces above (fxu) and below (fxd) the plane
      T fxn[CPL][4];
      if (hx) {
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
          if constexpr (down)
            face_ab<T>(nx[j], me[j], inv_dx, fxn[j]);
          else
            face_ab<T>(me[j], nx[j], inv_dx, fxn[j]);
        }
      }
      T(*fxu)[4] = down ? fxc : fxn;
      T(*fxd)[4] = down ? fxn : fxc;
      const unsigned r = row_base(x, gy);
#pragma unroll
      for (int j = 0; j < CPL; ++j) {
        // divergence in the reference order: x, y, z, each (F+ - F-) added in turn
        T acc[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          bool first = true;
          if (hx) {
            acc[a] = fxu[j][a] - fxd[j][a];
            first = false;
          }
          if (hy) {
            acc[a] = first ? (fyu[j][a] - fyd[j][a]) : ((acc[a] + fyu[j][a]) - fyd[j][a]);
            first = false;
          }
          acc[a] = first ? (fzu[j][a] - fzd[j][a]) : ((acc[a] + fzu[j][a]) - fzd[j][a]);
        }
        V4<T> o;
#pragma unroll
        for (int a = 0; a < 4; ++a) o.v[a] = acc[a] * dscale;
        if (kon[j]) {
#pragma unroll
          for (int a = 0; a < 4; ++a) tmax = max_ignore_nan<T>(tmax, fabs(o.v[a]));
          gst_field_o<T, V4<T>>(sd.dX, (r + kz[j]) * (unsigned)sizeof(V4<T>), o);
        }
      }
      if (hx) {
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
#pragma unroll
          for (int a = 0; a < 4; ++a) fxc[j][a] = fxn[j][a];
          me[j] = nx[j];
        }
      }
    }
  }
  if (cur >= 0) flush(cur);
  if (ndt > 0) {
    __shared__ int last;
    __shared__ double DRED[NW][NSTAT];
    if (tid == 0) {
      // the count depends on the values the maxima's atomics returned, so it is
      // made after they were performed (at device scope, where the controller's
      // exchange reads them); no cache write-back is needed for that.  The empty
      // asm hides folded's value from the compiler: without it, the compare below
      // (never true, since a stored maximum is a non-negative value whose bits
      // are not all ones) could be folded away, the maxima's atomics would lose
      // their returns and nothing would order them before the count.
      asm volatile("" : "+v"(folded));
      const unsigned one = 1u + (folded == ~0ull ? 1u : 0u);
      last = atomicAdd(done, one) == gridDim.x - 1;
    }
    __syncthreads();
    if (last) {
      for (int s = 0; s < ndt; ++s)
        dt_control<T, NT>(dt_list ? dt_list[s] : s, sims, st, xmax_acc, part, part_range, 1, DRED, tid);
      if (tid == 0) atomicExch(done, 0u);
    }
  }
}


// X of a lane's cells, packed for the LDS exchange of one plane's X values
template <typename T, int CPL, int QX>
PF5D_DEV void put_x(Q16<T> (*dst)[64], int lane, const V4<T> (&x)[CPL]) {
  constexpr int QN = 16 / sizeof(T);
  T p[QX * QN];
#pragma unroll
  for (int j = 0; j < CPL; ++j)
#pragma unroll
    for (int a = 0; a < 4; ++a) p[4 * j + a] = x[j].v[a];
#pragma unroll
  for (int i = 4 * CPL; i < QX * QN; ++i) p[i] = T(0);
#pragma unroll
  for (int q = 0; q < QX; ++q) {
    Q16<T> v;
#pragma unroll
    for (int e = 0; e < QN; ++e) v.v[e] = p[q * QN + e];
    sst<Q16<T>>(&dst[q][lane], v);
  }
}
template <typename T, int CPL, int QX>
PF5D_DEV void get_x(const Q16<T> (*src)[64], int lane, V4<T> (&x)[CPL]) {
  constexpr int QN = 16 / sizeof(T);
  T p[QX * QN];
#pragma unroll
  for (int q = 0; q < QX; ++q) {
    const Q16<T> v = sld<Q16<T>>(&src[q][lane]);
#pragma unroll
    for (int e = 0; e < QN; ++e) p[q * QN + e] = v.v[e];
  }
#pragma unroll
  for (int j = 0; j < CPL; ++j)
#pragma unroll
    for (int a = 0; a < 4; ++a) x[j].v[a] = p[4 * j + a];
}

// Fused marching stencil stage: interpolation, kappa lap(X), mu, face fluxes
// and divergence in one pass, for 3D fields whose rows (axis 2) hold 128
// cells (two per lane of a wavefront).
//
// An item is a tile of NW rows (one per wavefront) over a run of planes
// [xa, xb) on axis 0, stepped one plane at a time as in k_fluxmarch: the x
// face below a plane is carried from the previous plane, the y faces are
// exchanged between the wavefronts in LDS, the z faces between lanes.  The
// values a plane's faces need -- X, mu and Mob of its cells and of the next
// plane's -- come from two places:
// - the tile's inner rows (1 .. NW-2) on the run's inner planes (xa+1 ..
//   xb-2) are computed here, one plane ahead: interpolation from the item's
//   staged node box, the Laplacian of X with its y neighbours from the other
//   wavefronts (LDS, one plane ahead) and its z neighbours from the adjacent
//   lanes, and psi;
// - every other cell the tile reads -- its first and last rows, the run's first
//   and last planes, the plane below the run and above it, the rows beside the
//   tile -- lies on a boundary row (index 0 or NW-1 modulo NW) or a boundary
//   plane (xa - 1 or xa modulo the run length) of the simulation, whose mu and
//   Mob the potentials pass wrote (its work list holds only those cells).
// Each cell's mu and Mob are computed once, by one of the two kernels, with
// the same arithmetic (cell_mu, interp9_box4), so the result has the bits of
// the two-pass form; only the boundary cells' mu and Mob pass through memory.
// Items must hold NW rows of 128 cells and at least three planes; the host
// checks the shape.  The dt controller runs in the last block as in
// k_fluxmarch.
template <typename T, int NT, int MINW>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(MINW, 8)))
k_stencilmarch(const SimDev<T>* __restrict__ sims, SimState* __restrict__ st, T* __restrict__ xmax_acc,
               const K1Item* __restrict__ items, int nitems, const double* __restrict__ part,
               const int* __restrict__ part_range, int ndt, const int* __restrict__ dt_list,
               unsigned* __restrict__ done) {
  constexpr int CPL = 2;
  constexpr int NW = NT / 64;
  constexpr int QN = 16 / sizeof(T);
  constexpr int QC = (13 * CPL + QN - 1) / QN;  // 16-byte words per lane: cells
  constexpr int QF = (4 * CPL + QN - 1) / QN;   // 16-byte words per lane: one face, or X, per cell
  using L = BoxLayout<T>;
  static_assert(NW >= 3, "a tile needs inner rows");
  __shared__ __align__(16) T BOX[L::SLOTS * Vec16<T>::n];
  __shared__ Q16<T> SC[NW][QC][64];
  __shared__ Q16<T> SF[NW][QF][64];
  __shared__ Q16<T> SX[2][NW][QF][64];  // X of one plane, every row (by plane parity)
  __shared__ T RED[NW];
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));
  const int w = tid >> 6, lane = tid & 63;
  int ib, ie;
  block_range(nitems, ib, ie);
  T tmax = T(0);
  int cur = -1;
  unsigned long long folded = 0;  // previous bits returned by the maxima's atomics
  auto flush = [&](int sim) {
    const T m = wave_max<T>(tmax);
    if (lane == 0) RED[w] = m;
    __syncthreads();
    if (tid == 0) {
      T r = RED[0];
      for (int v = 1; v < NW; ++v) r = (RED[v] > r) ? RED[v] : r;
      folded |= atomic_max_nonneg(&xmax_acc[sim], r);
    }
    __syncthreads();
    tmax = T(0);
  };
  for (int it = ib; it < ie; ++it) {
    const K1Item item = items[it];
    if (item.sim != cur) {
      if (cur >= 0) flush(cur);
      cur = item.sim;
    }
    if (!(st[item.sim].status & ST_ACTIVE)) continue;  // uniform per block
    const SimDev<T>& sd = sims[item.sim];
    const int n0 = sd.n0, n1 = sd.n1, n2 = sd.n2;
    const T inv_dx = sd.inv_dx, dscale = step_div_scale<T>(inv_dx);
    const int xa = item.xa, xb = item.xb, run = xb - xa;
    const int gy = item.y0 + w;
    const bool inner = w > 0 && w < NW - 1;  // rows whose potentials are computed here
    auto xw = [&](int x) { return x < 0 ? x + n0 : (x >= n0 ? x - n0 : x); };
    // cell indices in 32 bits (the engine admits at most 2^32 / (4 sizeof(T)) cells)
    auto row_base = [&](int x, int y) { return ((unsigned)x * n1 + y) * n2 + CPL * lane; };
    auto load_full = [&](int x, int y, CellV<T>(&c)[CPL]) {
      const unsigned r = row_base(x, y);
#pragma unroll
      for (int j = 0; j < CPL; ++j) c[j] = load_cell32<T>(sd, r + j);
    };
    auto load_x = [&](int x, V4<T>(&v)[CPL]) {
      const unsigned r = row_base(x, gy);
#pragma unroll
      for (int j = 0; j < CPL; ++j) v[j] = gldo<V4<T>>(sd.X, (r + j) * (unsigned)sizeof(V4<T>));
    };
    // the row's psi at plane x, for the inner rows' potentials
    auto load_psi = [&](int x, T(&p)[CPL]) {
      const unsigned r = row_base(x, gy);
#pragma unroll
      for (int j = 0; j < CPL; ++j) p[j] = gldo<T>(sd.psi, (r + j) * (unsigned)sizeof(T));
    };
    int lo[4];
    stage_box4<T, NT>(sd, item, BOX, lo, tid);
    CellV<T> me[CPL];
    T fxd[CPL][4];
    load_full(xa, gy, me);
    {
      CellV<T> pv[CPL];
      load_full(xw(xa - 1), gy, pv);
#pragma unroll
      for (int j = 0; j < CPL; ++j) face_ab<T>(pv[j], me[j], inv_dx, fxd[j]);
    }
    // X of the next two planes (the next plane's cells and its x neighbour
    // above), and psi of the next plane
    V4<T> xn[CPL], xnn[CPL];
    T ps[CPL] = {};
    load_x(xw(xa + 1), xn);
    load_x(xw(xa + 2), xnn);
    if (inner && sd.elastic && run > 2) load_psi(xw(xa + 1), ps);
    put_x<T, CPL, QF>(SX[(xa + 1) & 1][w], lane, xn);
    __syncthreads();  // the node box and the next plane's X are in LDS
#pragma unroll 1
    for (int x = xa; x < xb; ++x) {
      const int r1 = x + 1 - xa;  // the next plane's place in the run
      const bool comp = inner && r1 <= run - 2;
      // loads for the planes ahead: X of plane x + 3, psi of plane x + 2
      V4<T> x3[CPL];
      T ps2[CPL] = {};
      if (x + 3 <= xb) load_x(xw(x + 3), x3);
      if (inner && sd.elastic && r1 + 1 <= run - 2) load_psi(xw(x + 2), ps2);
      CellV<T> nx[CPL];
      CellV<T> hal[CPL];
#pragma unroll
      for (int j = 0; j < CPL; ++j) nx[j].X = xn[j];
      if (comp) {
        // the next plane's mu and Mob: its X with the y neighbours from the
        // other rows, the z neighbours from the adjacent lanes (a row fills the
        // wavefront: periodic), the x neighbours from this plane and the one after
        V4<T> yu[CPL], yd[CPL], zu[CPL], zd[CPL];
        get_x<T, CPL, QF>(SX[(x + 1) & 1][w + 1], lane, yu);
        get_x<T, CPL, QF>(SX[(x + 1) & 1][w - 1], lane, yd);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          zu[0].v[a] = xn[1].v[a];
          zd[0].v[a] = rot_prev<T>(xn[1].v[a]);
          zu[1].v[a] = rot_next<T>(xn[0].v[a]);
          zd[1].v[a] = xn[0].v[a];
        }
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
          T P[9];
          interp9_box4<T, false>(sd.tab, BOX, lo, xn[j].v, P);
          T acc[4];
#pragma unroll
          for (int a = 0; a < 4; ++a) {
            acc[a] = xnn[j].v[a] + me[j].X.v[a];
            acc[a] = (acc[a] + yu[j].v[a]) + yd[j].v[a];
            acc[a] = (acc[a] + zu[j].v[a]) + zd[j].v[a];
          }
          nx[j].mu = cell_mu<T>(sd, xn[j], acc, P, ps[j]);
#pragma unroll
          for (int m = 0; m < 5; ++m) nx[j].M[m] = P[4 + m];
        }
      } else {
        // boundary rows and planes: the potentials pass wrote them
        const unsigned r = row_base(xw(x + 1), gy);
#pragma unroll
        for (int j = 0; j < CPL; ++j) {
          const unsigned o4 = (r + j) * (unsigned)sizeof(V4<T>);
          nx[j].mu = gldo<V4<T>>(sd.mu, o4);
          const V4<T> m = gldo<V4<T>>(sd.mb, o4);
#pragma unroll
          for (int a = 0; a < 4; ++a) nx[j].M[a] = m.v[a];
          nx[j].M[4] = gldo<T>(sd.mb1, (r + j) * (unsigned)sizeof(T));
        }
        // the rows beside the tile, for the edge rows' y faces
        if (w == NW - 1)
          load_full(x, gy + 1 < n1 ? gy + 1 : 0, hal);
        else if (w == 0)
          load_full(x, gy > 0 ? gy - 1 : n1 - 1, hal);
      }
      // the X of plane x + 2 for the next plane's inner rows, and this plane's
      // cells for the y faces
      if (r1 + 1 <= run - 2) put_x<T, CPL, QF>(SX[(x + 2) & 1][w], lane, xnn);
      put_cells<T, CPL, QC>(SC[w], lane, me);
      __syncthreads();
      T fyu[CPL][4], fyd[CPL][4];
      {
        CellV<T> up[CPL];
        if (w + 1 < NW)
          get_cells<T, CPL, QC>(SC[w + 1], lane, up);
        else {
#pragma unroll
          for (int j = 0; j < CPL; ++j) up[j] = hal[j];
        }
#pragma unroll
        for (int j = 0; j < CPL; ++j) face_ab<T>(me[j], up[j], inv_dx, fyu[j]);
      }
      {
        T p[QF * QN];
#pragma unroll
        for (int j = 0; j < CPL; ++j)
#pragma unroll
          for (int a = 0; a < 4; ++a) p[4 * j + a] = fyu[j][a];
#pragma unroll
        for (int q = 0; q < QF; ++q) {
          Q16<T> v;
#pragma unroll
          for (int e = 0; e < QN; ++e) v.v[e] = p[q * QN + e];
          sst<Q16<T>>(&SF[w][q][lane], v);
        }
      }
      __syncthreads();
      if (w > 0) {
        T p[QF * QN];
#pragma unroll
        for (int q = 0; q < QF; ++q) {
          const Q16<T> v = sld<Q16<T>>(&SF[w - 1][q][lane]);
#pragma unroll
          for (int e = 0; e < QN; ++e) p[q * QN + e] = v.v[e];
        }
#pragma unroll
        for (int j = 0; j < CPL; ++j)
#pragma unroll
          for (int a = 0; a < 4; ++a) fyd[j][a] = p[4 * j + a];
      } else {
#pragma unroll
        for (int j = 0; j < CPL; ++j) face_ab<T>(hal[j], me[j], inv_dx, fyd[j]);
      }
      // z: the face above each cell (the row's last cell: its first, periodic),
      // then the face below from the cell before
      T fzu[CPL][4], fzd[CPL][4];
      {
        const CellV<T> nb = rot_next_cell<T>(me[0]);
        face_ab<T>(me[0], me[1], inv_dx, fzu[0]);
        face_ab<T>(me[1], nb, inv_dx, fzu[1]);
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          fzd[1][a] = fzu[0][a];
          fzd[0][a] = rot_prev<T>(fzu[1][a]);
        }
      }
      T fxu[CPL][4];
#pragma unroll
      for (int j = 0; j < CPL; ++j) face_ab<T>(me[j], nx[j], inv_dx, fxu[j]);
      const unsigned r = row_base(x, gy);
#pragma unroll
      for (int j = 0; j < CPL; ++j) {
        // divergence in the reference order: x, y, z, each (F+ - F-) added in turn
        V4<T> o;
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          T acc = fxu[j][a] - fxd[j][a];
          acc = (acc + fyu[j][a]) - fyd[j][a];
          acc = (acc + fzu[j][a]) - fzd[j][a];
          o.v[a] = acc * dscale;
          tmax = max_ignore_nan<T>(tmax, fabs(o.v[a]));
        }
        gst_field_o<T, V4<T>>(sd.dX, (r + j) * (unsigned)sizeof(V4<T>), o);
      }
#pragma unroll
      for (int j = 0; j < CPL; ++j) {
#pragma unroll
        for (int a = 0; a < 4; ++a) fxd[j][a] = fxu[j][a];
        me[j] = nx[j];
        xn[j] = xnn[j];
        xnn[j] = x3[j];
        ps[j] = ps2[j];
      }
    }
    __syncthreads();  // BOX, SX are reused by the next item
  }
  if (cur >= 0) flush(cur);
  if (ndt > 0) {
    __shared__ int last;
    __shared__ double DRED[NW][NSTAT];
    if (tid == 0) {
      // as in k_fluxmarch: the count is made after the maxima's atomics, whose
      // returns it depends on (hidden from the compiler by the empty asm)
      asm volatile("" : "+v"(folded));
      const unsigned one = 1u + (folded == ~0ull ? 1u : 0u);
      last = atomicAdd(done, one) == gridDim.x - 1;
    }
    __syncthreads();
    if (last) {
      for (int s = 0; s < ndt; ++s)
        dt_control<T, NT>(dt_list ? dt_list[s] : s, sims, st, xmax_acc, part, part_range, 1, DRED, tid);
      if (tid == 0) atomicExch(done, 0u);
    }
  }
}

}  // namespace pf5d
