// Two-pass form of the stencil stage: the same arithmetic as k_stencil,
// operation for operation, so both forms give the same bits.
//
//   KP  k_potentials : per cell: interpolation + kappa lap(X) + elastic psi
//                      -> mu (4 values) and Mob (5 values) in HBM; the table
//                      nodes around the tile's first cell (4^4) are staged in
//                      LDS and serve every point whose cube lies among them
//   KF  k_fluxdiv    : per cell: the face fluxes to its neighbours from
//                      (X, mu, Mob) -> dX/dt, max |dX/dt| per simulation
//
// The single-pass form keeps mu and Mob in LDS but recomputes them on a halo
// and synchronises the block several times per plane; this form moves 9
// values per cell through HBM twice and needs one block barrier per tile (the
// node box), so many more waves stay resident.  Which is faster depends on
// the shape mix; the autotuner times both.
//
// Work items are (sim, x plane, y0, z0, ty, tz) tiles of at most NT cells:
// thread t handles cell (x, y0 + t / tz, z0 + t % tz).  Items are listed with
// x fastest and every block takes a contiguous range of them, so a block walks
// consecutive planes of one tile and finds its x neighbours in cache.  A face flux is
// evaluated by both cells it separates, always with the lower cell first, so
// both see the same value (the divergence telescopes exactly, as in the
// single-pass form where each face is evaluated once).
#pragma once
#include "pf5d_common.h"
#include "pf5d_flux.h"
#include "pf5d_fused.h"
#include "pf5d_interp.h"

namespace pf5d {

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

// MINW: minimum waves per SIMD the register allocation must allow
template <typename T, int NT, int MINW>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(MINW, 8)))
k_potentials(const SimDev<T>* __restrict__ sims, const SimState* __restrict__ st, const K1Item* __restrict__ items,
             int nitems) {
  constexpr int NP = TablePad<T>::value;
  constexpr int NBOX = BOX4 * BOX4 * BOX4 * BOX4;
  using VT = typename Vec16<T>::type;
  constexpr int NCOL = NP / Vec16<T>::n;  // 16-byte columns per table row
  __shared__ __align__(16) T BOX[NBOX * NP];
  int ib, ie;
  block_range(nitems, ib, ie);
  for (int it = ib; it < ie; ++it) {
    const K1Item item = items[it];
    if (!(st[item.sim].status & ST_ACTIVE)) continue;  // uniform per block
    // an empty asm keeps address arithmetic from being hoisted out of the item
    // loop (holding it across items costs more registers than recomputing it)
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const SimDev<T>& sd = sims[item.sim];
    const int n0 = sd.n0, n1 = sd.n1, n2 = sd.n2;
    const bool hx = sd.dim == 3, hy = sd.dim >= 2;
    int gx, gy, gz;
    const bool valid = tile_cell<T>(sd, item, tid, 1.0f / (float)item.tz, gx, gy, gz);
    const long c = ((long)gx * n1 + gy) * n2 + gz;
    V4<T> xc{};
    if (valid) xc = ld4g<T>(sd.X + 4 * c);
    // node box around the tile's first cell (the anchor): lower nodes one below its
    // lower cube corners, so corners within +-1 of the anchor's are covered
    const int num = sd.tab.num;
    int lo[4];
    {
      const V4<T> xa = ld4g<T>(sd.X + 4 * (((long)item.xa * n1 + item.y0) * n2 + item.z0));
#pragma unroll
      for (int a = 0; a < 4; ++a) lo[a] = min(max(floor_index<T>(xa.v[a] * sd.tab.inv_step, num) - 1, 0), num - BOX4);
    }
    for (int e = tid; e < NBOX * NCOL; e += NT) {
      const int node = e / NCOL, col = e - node * NCOL;
      const long row = ((((long)(lo[0] + (node >> 6)) * num + lo[1] + ((node >> 4) & 3)) * num + lo[2] +
                        ((node >> 2) & 3)) * num + lo[3] + (node & 3));
      reinterpret_cast<VT*>(BOX)[e] = gld<VT>(sd.tab.data + row * NP + col * Vec16<T>::n);
    }
    __syncthreads();
    if (valid) {
      T P[9];
      interp9_box4<T>(sd.tab, BOX, lo, xc.v, P);
      // neighbour sums in the reference order: x, then y, then z (each up + down)
      T acc[4];
      bool first = true;
      if (hx) {
        const long s = (long)n1 * n2;
        const V4<T> u = ld4g<T>(sd.X + 4 * (c + (gx + 1 < n0 ? s : s - (long)n0 * s)));
        const V4<T> d = ld4g<T>(sd.X + 4 * (c + (gx > 0 ? -s : (long)(n0 - 1) * s)));
#pragma unroll
        for (int a = 0; a < 4; ++a) acc[a] = u.v[a] + d.v[a];
        first = false;
      }
      if (hy) {
        const long u_i = c + (long)(wrap(gy + 1, n1) - gy) * n2, d_i = c + (long)(wrap(gy - 1, n1) - gy) * n2;
        const V4<T> u = ld4g<T>(sd.X + 4 * u_i), d = ld4g<T>(sd.X + 4 * d_i);
#pragma unroll
        for (int a = 0; a < 4; ++a) acc[a] = first ? (u.v[a] + d.v[a]) : ((acc[a] + u.v[a]) + d.v[a]);
        first = false;
      }
      {
        const V4<T> u = ld4g<T>(sd.X + 4 * (c + (wrap(gz + 1, n2) - gz)));
        const V4<T> d = ld4g<T>(sd.X + 4 * (c + (wrap(gz - 1, n2) - gz)));
#pragma unroll
        for (int a = 0; a < 4; ++a) acc[a] = first ? (u.v[a] + d.v[a]) : ((acc[a] + u.v[a]) + d.v[a]);
      }
      const T psi = sd.elastic ? gld<T>(sd.psi + c) : T(0);
      const T cdim = T(2 * sd.dim);
      V4<T> mu, mb;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        // (acc - 2 dim x) / h^2, then P - L kappa (+ psi l): simulate.py:253-255
        const T L = mad<T>(-cdim, xc.v[a], acc[a]) * sd.inv_dx2;
        T m = mad<T>(-L, sd.kappa[a], P[a]);
        if (sd.elastic) m = mad<T>(psi, sd.l[a], m);
        mu.v[a] = m;
        mb.v[a] = P[4 + a];
      }
      st4g<T>(sd.mu + 4 * c, mu);
      st4g<T>(sd.mb + 4 * c, mb);
      gst<T>(sd.mb1 + c, P[8]);
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

// flux through the face from cell a (lower) to cell b (upper)
template <typename T>
PF5D_DEV void face_ab(const CellV<T>& a, const CellV<T>& b, T inv_dx, T out[4]) {
  T g[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) g[k] = (b.mu.v[k] - a.mu.v[k]) * inv_dx;
  face_flux<T>(a.X.v, b.X.v, a.M, b.M, g, out);
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
    const bool valid = tile_cell<T>(sd, item, tid, 1.0f / (float)item.tz, gx, gy, gz);
    if constexpr (LOADS < 2) {
      if (!valid) continue;
    } else if (!valid) {
      // lanes past the tile's rows take part in the lane exchange on the tile's first cell
      gy = item.y0;
      gz = item.z0;
    }
    const long c = ((long)gx * n1 + gy) * n2 + gz;
    const T inv_dx = sd.inv_dx;
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
    for (int a = 0; a < 4; ++a) o.v[a] = acc[a] * inv_dx;
    if (valid) {
#pragma unroll
      for (int a = 0; a < 4; ++a) tmax = max_ignore_nan<T>(tmax, fabs(o.v[a]));
      st4g<T>(sd.dX + 4 * c, o);
    }
  }
  if (cur >= 0) flush(cur);
}

}  // namespace pf5d
