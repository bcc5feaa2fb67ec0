/* C/OpenMP FP64 restatement of the phasefield5d step loops -- TEST INFRASTRUCTURE
 * (CPU timing baseline for bench.py and a second CPU implementation for tests).
 *
 * Loop structure follows the reference's Numba kernels: `prange` over the
 * outermost spatial axis becomes `omp parallel for` over that axis; the
 * NumPy in-place passes of examples/simulate.py (L *= kappa, mu -= L,
 * mu += psi l, dX *= dt, X += dX) and abs_max stay single-threaded, as there.
 * Operation order per cell matches oracle/pf5d_oracle.py.
 *   interpolation   solver/interpolation.py:67-147
 *   laplacian       solver/operators.py:33-81
 *   face flux/div   solver/fluxes.py:128-160, 296-395
 *   misfit          solver/elastic.py:21-29
 *   abs_max, clamp  solver/operators.py:161-175, 262-317
 */
#include <math.h>
#include <omp.h>
#include <stddef.h>

static void axis_index(double u, int num, int* i, double* t) {
  if (u <= 0.0) { *i = 0; *t = 0.0; }
  else if (u >= num - 1) { *i = num - 2; *t = 1.0; }
  else { *i = (int)u; *t = u - (double)(*i); }
}

/* returns the number of cells outside the linear region (their rows are NaN) */
long ora_interp(const double* X, long ncell, const double* grid, int num, int np, double step, double* out) {
  const double thr = 1.0 - 4.0 * step - 1e-9, inv = 1.0 / step;
  const long n1 = num, n2 = (long)num * num, n3 = n2 * num;
  long nbad = 0;
#pragma omp parallel for reduction(+ : nbad) schedule(static)
  for (long c = 0; c < ncell; ++c) {
    const double* x = X + 4 * c;
    double* o = out + (long)np * c;
    if (!(((x[0] + x[1]) + x[2]) + x[3] <= thr)) {
      for (int p = 0; p < np; ++p) o[p] = NAN;
      nbad++;
      continue;
    }
    int i[4];
    double t[4], s[4];
    for (int d = 0; d < 4; ++d) {
      axis_index(x[d] * inv, num, &i[d], &t[d]);
      s[d] = 1.0 - t[d];
    }
    double w[16];
    long off[16];
    for (int b = 0; b < 16; ++b) {
      int b0 = (b >> 3) & 1, b1 = (b >> 2) & 1, b2 = (b >> 1) & 1, b3 = b & 1;
      double v = (b0 ? t[0] : s[0]) * (b1 ? t[1] : s[1]);
      v = v * (b2 ? t[2] : s[2]);
      v = v * (b3 ? t[3] : s[3]);
      w[b] = v;
      off[b] = ((((long)i[0] + b0) * n3 + ((long)i[1] + b1) * n2 + ((long)i[2] + b2) * n1 + (i[3] + b3))) * np;
    }
    for (int p = 0; p < np; ++p) {
      double acc = w[0] * grid[off[0] + p] + w[1] * grid[off[1] + p];
      for (int b = 2; b < 16; ++b) acc = acc + w[b] * grid[off[b] + p];
      o[p] = acc;
    }
  }
  return nbad;
}

static long idx3(const long* n, long i, long j, long k) { return (i * n[1] + j) * n[2] + k; }
static long wrapl(long i, long n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

/* shape padded to 3 axes (n0, n1, n2) with dim real axes first; neighbours in axis order */
void ora_laplacian(const double* X, int dim, const long* shp, double h, double* out) {
  long n[3] = {1, 1, 1};
  for (int a = 0; a < dim; ++a) n[a] = shp[a];
  const double inv2 = 1.0 / (h * h), cdim = 2.0 * dim;
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n[0]; ++i)
    for (long j = 0; j < n[1]; ++j)
      for (long k = 0; k < n[2]; ++k) {
        const long c = idx3(n, i, j, k);
        long nb[6];
        int m = 0;
        long q[3] = {i, j, k};
        for (int a = 0; a < dim; ++a)
          for (int st = 1; st >= -1; st -= 2) {
            long r[3] = {q[0], q[1], q[2]};
            r[a] = wrapl(r[a] + st, n[a]);
            nb[m++] = idx3(n, r[0], r[1], r[2]);
          }
        for (int s = 0; s < 4; ++s) {
          double acc = X[4 * nb[0] + s] + X[4 * nb[1] + s];
          for (int e = 2; e < m; ++e) acc = acc + X[4 * nb[e] + s];
          out[4 * c + s] = (acc - cdim * X[4 * c + s]) * inv2;
        }
      }
}

/* L *= kappa; mu(P[:, :4]) -= L  (single-threaded NumPy passes in the reference) */
void ora_kappa_sub(double* L, double* P, long ncell, const double* kappa) {
  for (long c = 0; c < ncell; ++c)
    for (int s = 0; s < 4; ++s) L[4 * c + s] = L[4 * c + s] * kappa[s];
  for (long c = 0; c < ncell; ++c)
    for (int s = 0; s < 4; ++s) P[9 * c + s] = P[9 * c + s] - L[4 * c + s];
}

void ora_misfit(const double* X, long ncell, const double* X0, const double* l, double* s) {
#pragma omp parallel for schedule(static)
  for (long c = 0; c < ncell; ++c) {
    double acc = 0.0;
    for (int a = 0; a < 4; ++a) acc += (X[4 * c + a] - X0[a]) * l[a];
    s[c] = acc;
  }
}

/* mu_elastic = psi (x) l ; mu += mu_elastic  (two NumPy passes) */
void ora_add_psi(double* P, const double* psi, long ncell, const double* l, double* tmp) {
  for (long c = 0; c < ncell; ++c)
    for (int s = 0; s < 4; ++s) tmp[4 * c + s] = psi[c] * l[s];
  for (long c = 0; c < ncell; ++c)
    for (int s = 0; s < 4; ++s) P[9 * c + s] = P[9 * c + s] + tmp[4 * c + s];
}

static void face_flux(const double* Xc, const double* Xn, const double* Mc, const double* Mn, const double* g,
                      double* out) {
  double Xm[4], Mm[5];
  for (int a = 0; a < 4; ++a) Xm[a] = 0.5 * (Xc[a] + Xn[a]);
  for (int m = 0; m < 5; ++m) Mm[m] = 0.5 * (Mc[m] + Mn[m]);
  double sumX = 0.0;
  for (int a = 0; a < 4; ++a) sumX += Xm[a];
  const double w_fe = (1.0 - sumX) * Mm[0];
  double dot = 0.0;
  for (int b = 0; b < 4; ++b) dot += Xm[b] * g[b];
  const double dot_fe = -dot;
  double dk[4];
  for (int k = 0; k < 4; ++k) dk[k] = g[k] - dot;
  double swk = 0.0;
  for (int k = 0; k < 4; ++k) swk += Xm[k] * Mm[1 + k] * dk[k];
  for (int a = 0; a < 4; ++a) {
    const double wa = Xm[a] * Mm[1 + a];
    out[a] = w_fe * (-Xm[a]) * dot_fe + wa * dk[a] - Xm[a] * swk;
  }
}

/* P rows of 9: mu = P[0:4], Mob = P[4:9] */
void ora_div_from_mu(const double* X, const double* P, int dim, const long* shp, double inv_dx, double* out) {
  long n[3] = {1, 1, 1};
  for (int a = 0; a < dim; ++a) n[a] = shp[a];
#pragma omp parallel for schedule(static)
  for (long i = 0; i < n[0]; ++i)
    for (long j = 0; j < n[1]; ++j)
      for (long k = 0; k < n[2]; ++k) {
        const long c = idx3(n, i, j, k);
        long q[3] = {i, j, k};
        double acc[4] = {0, 0, 0, 0};
        for (int a = 0; a < dim; ++a) {
          double f[2][4];
          for (int e = 0; e < 2; ++e) {
            const int st = e == 0 ? 1 : -1;
            long r[3] = {q[0], q[1], q[2]};
            r[a] = wrapl(r[a] + st, n[a]);
            const long nb = idx3(n, r[0], r[1], r[2]);
            double g[4];
            for (int s = 0; s < 4; ++s)
              g[s] = (st > 0 ? (P[9 * nb + s] - P[9 * c + s]) : (P[9 * c + s] - P[9 * nb + s])) * inv_dx;
            face_flux(X + 4 * c, X + 4 * nb, P + 9 * c + 4, P + 9 * nb + 4, g, f[e]);
          }
          for (int s = 0; s < 4; ++s) acc[s] = (a == 0) ? (f[0][s] - f[1][s]) : ((acc[s] + f[0][s]) - f[1][s]);
        }
        for (int s = 0; s < 4; ++s) out[4 * c + s] = acc[s] * inv_dx;
      }
}

double ora_abs_max(const double* a, long n) {
  double m = 0.0;
  for (long i = 0; i < n; ++i) {
    const double v = fabs(a[i]);
    if (v > m) m = v;
  }
  return m;
}

void ora_update(double* X, double* dX, long n, double dt) {
  for (long i = 0; i < n; ++i) dX[i] = dX[i] * dt;
  for (long i = 0; i < n; ++i) X[i] = X[i] + dX[i];
}

void ora_clamp(double* X, long ncell) {
#pragma omp parallel for schedule(static)
  for (long c = 0; c < ncell; ++c) {
    double tot = 0.0;
    for (int s = 0; s < 4; ++s) tot += X[4 * c + s];
    if (tot > 1.0) {
      const double inv = 1.0 / tot;
      for (int s = 0; s < 4; ++s) X[4 * c + s] *= inv;
    }
  }
}

int ora_threads(void) { return omp_get_max_threads(); }
void ora_set_threads(int n) { omp_set_num_threads(n); }
