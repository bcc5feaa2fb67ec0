"""NumPy FP64 restatement of the phasefield5d explicit step (test oracle).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py.  Never part of the product path.

Each function cites the reference code it restates and keeps its floating
point operation order (left-to-right sums, no FMA contraction), so the device
FP64 path can be compared bit-for-bit where no FFT is involved.

Pinning of each function (tests/test_oracle_pins.py):
  interpolate_grid_linear  exact on multilinear f = sum_S c_S prod X_d (any node/
                           interior point); scipy RegularGridInterpolator(linear)
                           on interior points; closed-form edge clamps.
  nearest fallback         scipy cKDTree (the reference's own dependency) and a
                           brute-force scan on small tables.
  laplacian                Fourier eigenvalues sum_d (2cos(2 pi m/N) - 2)/h^2;
                           periodic sum-zero invariant (reference test_smoke.py:24-40).
  face_flux                equals the dyadic mobility matrix product M g built
                           independently (reference solver/mobility.py:36-78);
                           constant Mob -> (D/RT)(diag X - X X^T) g.
  divergence_from_mu       uniform X + Fourier-mode mu -> M(X0) lap_h mu exactly;
                           sum over cells = 0 (telescoping).
  elastic_psi              1D: B constant for k != 0 -> psi = B (s - mean s);
                           plane wave -> B(k) cos(k.x); B(0) = 0.
  misfit_scalar /          X = X0 + a v cos(k.x) -> mu += B(k) a (l.v) cos(k.x) l
  elastic_update           (2D in-plane and 3D); the misfit is linear with slope l_j.
  abs_max / adaptive dt /  closed-form scalar rules on an exhaustive input grid;
  clamp / out_of_bounds    each hard bounds branch alone, soft-only cases.
  apply_fluctuations       a hand-built draw: the seeded legacy generator's doubles
                           one at a time, five per cell, mean over all five removed.
"""
import numpy as np
import scipy.fft as sfft
from scipy.spatial import cKDTree

N_COMP = 4
N_PROPS = 9


# ---------------------------------------------------------------------------
# table -> dense grid (solver/interpolation.py:13-44)
# ---------------------------------------------------------------------------
def build_4d_grid(rows, resolution):
    rows = np.asarray(rows, dtype=np.float64)
    num = int(round(1.0 / float(resolution))) + 1
    idx = np.clip(np.rint(rows[:, :4] / float(resolution)).astype(np.int32), 0, num - 1)
    grid = np.full((num,) * 4 + (rows.shape[1] - 4,), np.nan)
    grid[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3], :] = rows[:, 4:]
    return grid


def build_kdtree(rows):
    rows = np.asarray(rows, dtype=np.float64)
    return cKDTree(rows[:, :4]), rows[:, 4:].copy()


def linear_threshold(resolution):
    """solver/interpolation.py:195."""
    return 1.0 - 4.0 * resolution - 1e-9


# ---------------------------------------------------------------------------
# multilinear interpolation (solver/interpolation.py:67-147, 154-211)
# ---------------------------------------------------------------------------
def _axis_index(u, num):
    """interpolation.py:110-121: clamp-to-edge floor index and fraction."""
    lo = u <= 0.0
    hi = (~lo) & (u >= num - 1)
    mid = ~(lo | hi)
    i = np.zeros(u.shape, dtype=np.int64)
    t = np.zeros(u.shape, dtype=np.float64)
    i[hi] = num - 2
    t[hi] = 1.0
    im = u[mid].astype(np.int64)
    i[mid] = im
    t[mid] = u[mid] - im
    return i, t


def interpolate_grid_linear(Xf, grid, step):
    """(M,4) -> (M,P): 16-corner sum in the order c0000..c1111 (i3 fastest)."""
    num = grid.shape[0]
    inv_step = 1.0 / step
    idx, tt = [], []
    for d in range(4):
        i, t = _axis_index(Xf[:, d] * inv_step, num)
        idx.append(i)
        tt.append(t)
    ss = [1.0 - t for t in tt]
    out = None
    for b in range(16):
        bits = [(b >> (3 - d)) & 1 for d in range(4)]
        w = (tt[0] if bits[0] else ss[0]) * (tt[1] if bits[1] else ss[1])
        w = w * (tt[2] if bits[2] else ss[2])
        w = w * (tt[3] if bits[3] else ss[3])
        g = grid[idx[0] + bits[0], idx[1] + bits[1], idx[2] + bits[2], idx[3] + bits[3], :]
        term = w[:, None] * g
        out = term if out is None else out + term
    return out


def interpolator(X, grid, resolution, tree, values):
    """interpolator_nb: linear where sum(X) <= threshold, nearest table row elsewhere."""
    X = np.asarray(X, dtype=np.float64)
    Xf = X.reshape(-1, 4)
    s = ((Xf[:, 0] + Xf[:, 1]) + Xf[:, 2]) + Xf[:, 3]
    mask = s <= linear_threshold(resolution)
    out = np.full((Xf.shape[0], grid.shape[-1]), np.nan)
    if mask.any():
        out[mask] = interpolate_grid_linear(Xf[mask], grid, float(resolution))
    if (~mask).any():
        _, nn = tree.query(Xf[~mask])
        out[~mask] = values[nn]
    return out.reshape(X.shape[:-1] + (grid.shape[-1],))


# ---------------------------------------------------------------------------
# periodic Laplacian (solver/operators.py:33-81)
# ---------------------------------------------------------------------------
def _nbr(a, axis, step):
    """value at index i+step (periodic) along ``axis``."""
    return np.roll(a, -step, axis=axis)


def laplacian(X, h):
    dim = X.ndim - 1
    inv_dx2 = 1.0 / (h * h)
    acc = None
    for d in range(dim):
        for st in (1, -1):
            v = _nbr(X, d, st)
            acc = v if acc is None else acc + v
    return (acc - (2.0 * dim) * X) * inv_dx2


# ---------------------------------------------------------------------------
# face flux (solver/fluxes.py:128-160) and fused divergence (:296-395)
# ---------------------------------------------------------------------------
def face_flux(Xc, Xn, Mc, Mn, g):
    Xm = 0.5 * (Xc + Xn)
    Mm = 0.5 * (Mc + Mn)
    sumX = ((Xm[..., 0] + Xm[..., 1]) + Xm[..., 2]) + Xm[..., 3]
    w_fe = (1.0 - sumX) * Mm[..., 0]
    dot = ((Xm[..., 0] * g[..., 0] + Xm[..., 1] * g[..., 1]) + Xm[..., 2] * g[..., 2]) + Xm[..., 3] * g[..., 3]
    dot_fe = -dot
    dk = g - dot[..., None]
    wk = Xm * Mm[..., 1:]
    swk = ((wk[..., 0] * dk[..., 0] + wk[..., 1] * dk[..., 1]) + wk[..., 2] * dk[..., 2]) + wk[..., 3] * dk[..., 3]
    return (w_fe[..., None] * (-Xm)) * dot_fe[..., None] + wk * dk - Xm * swk[..., None]


def divergence_from_mu(X, Mob, mu, inv_dx):
    dim = X.ndim - 1
    acc = None
    for d in range(dim):
        Xp, Xm_ = _nbr(X, d, 1), _nbr(X, d, -1)
        Mp, Mm_ = _nbr(Mob, d, 1), _nbr(Mob, d, -1)
        mup, mum = _nbr(mu, d, 1), _nbr(mu, d, -1)
        gp = (mup - mu) * inv_dx
        gm = (mu - mum) * inv_dx
        fp = face_flux(X, Xp, Mob, Mp, gp)
        fm = face_flux(X, Xm_, Mob, Mm_, gm)
        acc = (fp - fm) if acc is None else (acc + fp) - fm
    return acc * inv_dx


def gradients_pm(A, h):
    """calculate_gradients_pm (solver/operators.py:88-158): (dim, *spatial, n) forward
    (A(c+e_d) - A(c)) / h and backward (A(c) - A(c-e_d)) / h differences."""
    dim = A.ndim - 1
    inv_dx = 1.0 / h
    gp = np.stack([(_nbr(A, d, 1) - A) * inv_dx for d in range(dim)])
    gm = np.stack([(A - _nbr(A, d, -1)) * inv_dx for d in range(dim)])
    return gp, gm


def fluxes_pm(X, Mob, gp, gm):
    """get_flux_function(dim) (solver/fluxes.py:6-125): face fluxes towards the +/- neighbour
    along every axis, (dim, *spatial, 4) each."""
    dim = X.ndim - 1
    fp = np.stack([face_flux(X, _nbr(X, d, 1), Mob, _nbr(Mob, d, 1), gp[d]) for d in range(dim)])
    fm = np.stack([face_flux(X, _nbr(X, d, -1), Mob, _nbr(Mob, d, -1), gm[d]) for d in range(dim)])
    return fp, fm


# ---------------------------------------------------------------------------
# rank-1 elastic term (solver/elastic.py:21-29, 246-263)
# ---------------------------------------------------------------------------
def misfit_scalar(X, X0, l):
    t = (X - X0) * l
    return ((t[..., 0] + t[..., 1]) + t[..., 2]) + t[..., 3]


def elastic_psi(s, Bhat):
    axes = tuple(range(s.ndim))
    k = sfft.rfftn(s, axes=axes, workers=1)
    k *= Bhat
    return sfft.irfftn(k, s=s.shape, axes=axes, workers=1).real


def elastic_update(mu, X, X0, l, Bhat):
    psi = elastic_psi(misfit_scalar(X, X0, l), Bhat)
    mu += psi[..., None] * l
    return mu


# ---------------------------------------------------------------------------
# scalar controller, update, clamp, bounds (operators.py:161-175, 262-317;
# system.py:12-53, 84-112; simulate.py:265-320)
# ---------------------------------------------------------------------------
def abs_max(a):
    a = np.abs(np.asarray(a, dtype=np.float64)).ravel()
    a = a[~np.isnan(a)]
    return float(a.max()) if a.size and a.max() > 0.0 else 0.0


def apply_adaptive_time(dt, xdotmax, ctld, t_cfl, lower=1e-5, upper=1e-3, factor=2.0):
    if xdotmax is None or xdotmax <= 0.0:
        raw = dt
    else:
        lo, hi = lower / xdotmax, upper / xdotmax
        if lo > hi:
            lo, hi = hi, lo
        raw = min(max(dt, lo), hi)
    return min(min(dt * factor, raw), 0.005 * ctld, t_cfl)


def clamp_simplex(X):
    tot = ((X[..., 0] + X[..., 1]) + X[..., 2]) + X[..., 3]
    over = tot > 1.0
    inv = 1.0 / tot[over]
    X[over] *= inv[:, None]
    return X


def out_of_bounds(X, tol_soft=1e-5, tol_hard=1e-3):
    if not np.isfinite(X).all():
        return True
    s = X.sum(axis=-1)
    return bool(s.max() > 1.0 + tol_hard or s.min() < -tol_hard
                or X.min() < -tol_hard or X.max() > 1.0 + tol_hard)


def apply_fluctuations(X0, shape, fluctuation, seed=None):
    """system.py:56-61 with an optional np.random.seed before the identical draw."""
    if seed is not None:
        np.random.seed(seed)
    X0 = np.asarray(X0, dtype=np.float64)
    noise = (np.random.rand(*shape, X0.shape[-1] + 1) - 0.5) * 2.0
    noise -= noise.mean(axis=-1, keepdims=True)
    return np.broadcast_to(X0, (*shape, X0.shape[-1])) + fluctuation * noise[..., 1:]


# ---------------------------------------------------------------------------
# one explicit step (simulate.py:242-320)
# ---------------------------------------------------------------------------
class OracleSim:
    """Constants of one simulation, in the units and layouts the step uses."""

    def __init__(self, grid, resolution, tree, values, h, kappa, X0, l_vec=None, Bhat=None,
                 ctld=1.0, dt_cfl=np.inf, lower=1e-5, upper=1e-3):
        self.grid, self.resolution, self.tree, self.values = grid, float(resolution), tree, values
        self.h = float(h)
        self.kappa = np.asarray(kappa, dtype=np.float64).reshape(1, 4)
        self.X0 = np.asarray(X0, dtype=np.float64)
        self.l_vec = None if l_vec is None else np.asarray(l_vec, dtype=np.float64)
        self.Bhat = Bhat
        self.ctld, self.dt_cfl, self.lower, self.upper = float(ctld), float(dt_cfl), lower, upper

    def chemical_potential(self, X):
        P = interpolator(X, self.grid, self.resolution, self.tree, self.values)
        mu = P[..., :4].copy()
        L = laplacian(X, self.h)
        L *= self.kappa
        mu -= L
        if self.l_vec is not None and self.Bhat is not None:
            elastic_update(mu, X, self.X0, self.l_vec, self.Bhat)
        return P, mu

    def rate(self, X):
        P, mu = self.chemical_potential(X)
        return divergence_from_mu(X, P[..., 4:], mu, 1.0 / self.h)

    def step(self, X, dt, t=0.0):
        """Return (X_next, dXdt, dt_new, t_next, xmax); X is not modified."""
        dX = self.rate(X)
        xmax = abs_max(dX)
        dt_new = apply_adaptive_time(dt, xmax, self.ctld, self.dt_cfl, self.lower, self.upper)
        Xn = X + dX * dt_new
        clamp_simplex(Xn)
        return Xn, dX, dt_new, t + dt_new, xmax
