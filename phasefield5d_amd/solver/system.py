"""System-level helpers (phasefield5d/solver/system.py).

Host setup: fluctuations, dt0, CFL bound, grid sizing, Fourier grids, initial
kinetics (CTLD).  ``apply_adaptive_time`` is the host form of the dt rule the
device controller also implements (k_dtctl).  ``out_of_bounds`` reduces on the
GPU.
"""
import ctypes as C

import numpy as np

from .. import _lib
from ..device import Staged, prec_of
from ..kinetics.operators import calculate_kinetic_operators, calculate_kinetic_eigenvalues_vectors
from ..utils.statistics import round_to_first_nonzero, round_down_to_first_nonzero


def bounds_statistics(composition):
    """[nonfinite count, max sum, min sum, min component, max component] (device reduction)."""
    prec = prec_of(composition)
    X = Staged(composition)
    Xd = X.dev.contiguous()
    out = (C.c_double * 5)()
    _lib.check(_lib.lib().pf5d_bounds_stats(prec, C.c_void_p(Xd.data_ptr()), Xd.numel() // 4, out,
                                           _lib.stream_handle()))
    return list(out)


def out_of_bounds(composition, tol_soft=1e-5, tol_hard=1e-3):
    """True only for hard violations; soft ones are printed (system.py:12-53)."""
    nonfinite, smax, smin, cmin, cmax = bounds_statistics(composition)
    if nonfinite > 0:
        print(f"HARD OUT OF BOUNDS: non-finite value (NaN/inf) in {int(nonfinite)} cells")
        return True
    soft, hard = [], []
    checks = [(smax > 1.0 + tol_soft, smax > 1.0 + tol_hard, f"sum(N-1) > 1 + tol. max sum={smax:.8f}"),
              (smin < -tol_soft, smin < -tol_hard, f"sum(N-1) < -tol. min sum={smin:.8f}"),
              (cmin < -tol_soft, cmin < -tol_hard, f"component < -tol. min comp={cmin:.8f}"),
              (cmax > 1.0 + tol_soft, cmax > 1.0 + tol_hard, f"component > 1 + tol. max comp={cmax:.8f}")]
    for s, h, msg in checks:
        if s:
            soft.append("SOFT OOB: " + msg)
        if h:
            hard.append("HARD OOB: " + msg)
    for m in soft:
        print(m)
    for m in hard:
        print(m)
    return bool(hard)


def apply_fluctuations(reference_composition, size, fluctuation, seed=None):
    """X0 + f * (uniform(-1,1) noise over 5 components, per-cell mean removed)[..., 1:] (system.py:56-61).

    ``seed`` (added) draws from a legacy MT19937 generator seeded with it: the same
    sequence as seeding NumPy's global generator first, without touching the global
    state, so concurrent set-ups (the server's worker threads) stay reproducible.
    Without a seed the global generator is used, as in the reference.  The
    arithmetic is the reference's, done in place.
    """
    rs = np.random.RandomState(seed) if seed is not None else np.random
    ref = np.asarray(reference_composition, dtype=np.float64)
    n = ref.shape[-1]
    noise = rs.rand(*size, n + 1)
    noise -= 0.5
    noise *= 2.0
    noise -= noise.mean(axis=-1, keepdims=True)
    out = noise[..., 1:] * fluctuation
    out += ref
    return out


def snap_to_grid_floor(comp, step=0.01):
    """system.py:64-73: rows of compositions (n, k) clipped to [0, 1], rows summing above
    1 scaled back onto the simplex, then floored to the table grid of spacing ``step``
    (with a 1e-9 guard so values already on a node stay there).  Host set-up helper."""
    c = np.clip(np.array(comp, dtype=float), 0.0, 1.0)
    tot = c.sum(axis=-1, keepdims=True)
    over = tot[:, 0] > 1.0
    c[over] = c[over] / tot[over]
    return np.floor(c * (1.0 / step) + 1e-9).astype(int) / (1.0 / step)


def get_time_increment(ctld, steps_per_ctld=100):
    return round_down_to_first_nonzero(ctld / steps_per_ctld)


def cahn_hilliard_cfl_dt(dx, mobility_max, hessian_max, kappa_max, safety=1):
    kmax = np.pi / dx
    return safety * (2.0 / (mobility_max * (kappa_max * kmax**4 + abs(hessian_max) * kmax**2)))


def apply_adaptive_time(dt_current, xdotmax, ctld, t_cfl, *, composition_change_lower_limit=1e-5,
                        composition_change_upper_limit=1e-3, max_change_factor=2.0):
    """min(2 dt, clip(dt, lo/x, hi/x), 0.005 ctld, t_cfl) (system.py:90-112)."""
    if xdotmax is None or xdotmax <= 0.0:
        raw = dt_current
    else:
        lo = composition_change_lower_limit / xdotmax
        hi = composition_change_upper_limit / xdotmax
        if lo > hi:
            lo, hi = hi, lo
        raw = float(np.clip(dt_current, lo, hi))
    out = min(dt_current * max_change_factor, raw)
    return min(out, 0.005 * ctld, t_cfl)


def get_system_dimensions(wavenumber_max, ppw=16, mw=100, dim=1):
    wavelength = 2 * np.pi / wavenumber_max
    return (wavelength, round_down_to_first_nonzero(wavelength / ppw),
            round_to_first_nonzero(mw * wavelength), (mw * ppw,) * dim)


def build_fourier_grid(number_of_cells, dx):
    ks = [2.0 * np.pi * np.fft.fftfreq(n, d=dx) for n in number_of_cells]
    k_grid = np.stack(np.meshgrid(*ks, indexing="ij"), axis=0)
    return k_grid, np.linalg.norm(k_grid, axis=0)


def build_fourier_grid_1d_along_direction(n, dx, direction):
    """rfft k-grid of a 1D run embedded along ``direction`` (system.py:140-153)."""
    d = np.asarray(direction, dtype=float)
    d = d / np.linalg.norm(d)
    q = 2.0 * np.pi * np.fft.rfftfreq(n, d=dx)
    return d[:, None] * q[None, :], np.abs(q)


def build_fourier_grid_2d_in_plane(nx, ny, dx, normal):
    """rfft k-grid of a 2D run in the plane normal to ``normal`` (system.py:156-190)."""
    nrm = np.asarray(normal, dtype=float)
    nrm = nrm / np.linalg.norm(nrm)
    cand = np.array([1.0, 0.0, 0.0])
    if abs(np.dot(cand, nrm)) > 0.9:
        cand = np.array([0.0, 1.0, 0.0])
    e1 = cand - np.dot(cand, nrm) * nrm
    e1 = e1 / np.linalg.norm(e1)
    e2 = np.cross(nrm, e1)
    e2 = e2 / np.linalg.norm(e2)
    kx = 2.0 * np.pi * np.fft.fftfreq(nx, d=dx)
    ky = 2.0 * np.pi * np.fft.rfftfreq(ny, d=dx)
    Kx, Ky = np.meshgrid(kx, ky, indexing="ij")
    k_grid = e1[:, None, None] * Kx[None] + e2[:, None, None] * Ky[None]
    return k_grid, np.sqrt(Kx**2 + Ky**2)


def build_fourier_grid_3d(nx, ny, nz, dx):
    kx = 2.0 * np.pi * np.fft.fftfreq(nx, d=dx)
    ky = 2.0 * np.pi * np.fft.fftfreq(ny, d=dx)
    kz = 2.0 * np.pi * np.fft.rfftfreq(nz, d=dx)
    k_grid = np.stack(np.meshgrid(kx, ky, kz, indexing="ij"), axis=0)
    return k_grid, np.linalg.norm(k_grid, axis=0)


def solve_elastic_kinetic_eigenproblem(hessian, mobility_matrix, kappa_matrix, n_k=10_000, verbose=False):
    ks, ops = calculate_kinetic_operators(hessian, mobility_matrix, kappa_matrix, n_k=n_k)
    rate, kmax, val, mode = calculate_kinetic_eigenvalues_vectors(ks, ops, max_tag=True)
    if verbose:
        print(f"  Max growth rate : {rate: .4e} [1/s]   Max wavenumber : {kmax: .4e} [1/m]")
    return rate, kmax


def dyadic_mobility_matrix(X0, mob5):
    """M = X_Fe M_Fe X X^T + sum_k X_k M_k (e_k - X)(e_k - X)^T at one composition (system.py:304-312)."""
    X0 = np.asarray(X0, dtype=float)
    n = X0.shape[0]
    v_sol = np.eye(n) - X0[None, :]
    return ((1.0 - X0.sum()) * mob5[0] * np.outer(-X0, -X0)
            + np.einsum("k,ka,kb->ab", X0 * mob5[1:], v_sol, v_sol))


def compute_initial_kinetics(initial_composition, resolution, kappa_value, kappa_i, data_grid, tree,
                             calphad_values, elastic_matrix=None, n_k=10_000, verbose=False):
    """(ctld, k_max) from the FD Hessian on the table and the dyadic mobility at X0 (system.py:241-329)."""
    from .interpolation import interpolate_points_host
    X0 = np.asarray(initial_composition, dtype=float)
    n = X0.shape[0]
    pts = np.vstack([X0[None, :], X0[None, :] + np.eye(n) * resolution, X0[None, :] - np.eye(n) * resolution])
    data = interpolate_points_host(pts, data_grid, resolution, tree, calphad_values)
    hessian = (data[1:n + 1, :n] - data[n + 1:, :n]).T / (2.0 * resolution)
    mob = dyadic_mobility_matrix(X0, data[0, n:])
    H = hessian if elastic_matrix is None else hessian + elastic_matrix
    kappa_mat = np.diag(np.asarray(kappa_i, dtype=float).flatten() * kappa_value)
    rate, kmax = solve_elastic_kinetic_eigenproblem(H, mob, kappa_mat, n_k=n_k, verbose=verbose)
    if not rate > 0.0:
        raise ValueError(f"No unstable mode found (max_growth_rate={rate:.3e}). "
                         "Composition may be outside the spinodal region.")
    return 1.0 / rate, kmax
