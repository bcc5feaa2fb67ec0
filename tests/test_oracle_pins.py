"""Pins of the FP64 oracle against closed forms, invariants and independent implementations.

Each check is chosen so that a plausible slip in the oracle (a dropped term, a
wrong sign or index, a transposed operand) fails it; none of them retypes the
oracle's own formula.
"""
import itertools

import numpy as np
import pytest
from scipy.interpolate import RegularGridInterpolator
from scipy.spatial import cKDTree

from oracle import pf5d_oracle as O

H = 0.025
NUM = 41


def multilinear_grid(rng, num=NUM, step=H, nprops=9):
    """Grid values of f_p(X) = sum_S c_{S,p} prod_{d in S} X_d (multilinear), and f itself."""
    subsets = [s for r in range(5) for s in itertools.combinations(range(4), r)]
    coef = rng.normal(size=(len(subsets), nprops))

    def f(X):
        X = np.atleast_2d(X)
        out = np.zeros((X.shape[0], nprops))
        for S, c in zip(subsets, coef):
            term = np.ones(X.shape[0])
            for d in S:
                term = term * X[:, d]
            out += term[:, None] * c[None, :]
        return out

    axes = np.arange(num) * step
    G = np.stack(np.meshgrid(axes, axes, axes, axes, indexing="ij"), axis=-1).reshape(-1, 4)
    grid = f(G).reshape((num,) * 4 + (nprops,))
    return grid, f


def interior_points(rng, n, smax=0.88):
    X = rng.dirichlet([2, 2, 2, 2, 2], size=n)[:, 1:]
    return X[X.sum(axis=1) <= smax]


def test_interp_exact_on_multilinear_functions():
    rng = np.random.default_rng(0)
    grid, f = multilinear_grid(rng)
    X = interior_points(rng, 400)
    got = O.interpolate_grid_linear(X, grid, H)
    assert np.allclose(got, f(X), rtol=1e-12, atol=1e-12)
    # on the nodes themselves
    nodes = rng.integers(0, 8, size=(50, 4)) * H
    assert np.allclose(O.interpolate_grid_linear(nodes, grid, H), f(nodes), rtol=1e-12, atol=1e-12)


def test_interp_matches_scipy_regular_grid_interpolator():
    rng = np.random.default_rng(1)
    grid = rng.normal(size=(NUM,) * 4 + (9,))
    ax = np.linspace(0.0, 1.0, NUM)
    rgi = RegularGridInterpolator((ax, ax, ax, ax), grid, method="linear")
    X = interior_points(rng, 300)
    assert np.allclose(O.interpolate_grid_linear(X, grid, H), rgi(X), rtol=1e-11, atol=1e-11)


def test_interp_edge_clamps():
    rng = np.random.default_rng(2)
    grid = rng.normal(size=(NUM,) * 4 + (9,))
    # u <= 0 on an axis behaves as X_d = 0 (index 0, weight 0)
    X = np.array([[-0.01, 0.2, 0.1, 0.3]])
    X0 = np.array([[0.0, 0.2, 0.1, 0.3]])
    assert np.array_equal(O.interpolate_grid_linear(X, grid, H), O.interpolate_grid_linear(X0, grid, H))
    # u >= num-1 uses the last cell with t = 1, i.e. the value on the last node
    Xh = np.array([[1.0, 0.0, 0.0, 0.0]])
    assert np.allclose(O.interpolate_grid_linear(Xh, grid, H)[0], grid[NUM - 1, 0, 0, 0], rtol=0, atol=0)


def test_nearest_fallback_brute_force():
    rng = np.random.default_rng(3)
    step = 0.25
    idx = np.array([i for i in itertools.product(range(5), repeat=4) if sum(i) <= 4])
    coords = idx * step
    vals = rng.normal(size=(len(idx), 9))
    rows = np.concatenate([coords, vals], axis=1)
    grid = O.build_4d_grid(rows, step)
    tree, values = O.build_kdtree(rows)
    X = rng.dirichlet([1, 1, 1, 1, 1], size=200)[:, 1:]
    X = X[X.sum(axis=1) > O.linear_threshold(step)]
    got = O.interpolator(X, grid, step, tree, values)
    d = ((X[:, None, :] - coords[None, :, :]) ** 2).sum(axis=-1)
    assert np.array_equal(got, vals[np.argmin(d, axis=1)])


@pytest.mark.parametrize("shape", [(32,), (16, 12), (8, 6, 10)])
def test_laplacian_fourier_eigenvalues_and_sum_zero(shape):
    h = 0.7
    rng = np.random.default_rng(4)
    m = [int(rng.integers(0, n)) for n in shape]
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    phase = sum(2 * np.pi * m[d] * grids[d] / shape[d] for d in range(len(shape)))
    X = np.stack([np.cos(phase + s) for s in range(4)], axis=-1)
    lam = sum((2 * np.cos(2 * np.pi * m[d] / shape[d]) - 2) / h**2 for d in range(len(shape)))
    assert np.allclose(O.laplacian(X, h), lam * X, atol=1e-12 * max(1, abs(lam)))
    Y = rng.random(shape + (4,))
    assert np.allclose(O.laplacian(Y, h).reshape(-1, 4).sum(axis=0), 0.0, atol=1e-10)


def dyadic_matrix(Xm, Mm):
    """M = w_Fe X X^T + sum_k X_k M_k (e_k - X)(e_k - X)^T, built independently (mobility.py:36-78)."""
    x_fe = 1.0 - Xm.sum()
    M = x_fe * Mm[0] * np.outer(Xm, Xm)
    for k in range(4):
        e = np.eye(4)[k] - Xm
        M += Xm[k] * Mm[1 + k] * np.outer(e, e)
    return M


def test_face_flux_is_dyadic_mobility_times_gradient():
    rng = np.random.default_rng(5)
    for _ in range(20):
        Xa, Xb = rng.dirichlet(np.ones(5), 2)[:, 1:]
        Ma, Mb = rng.uniform(0.5, 2, (2, 5))
        g = rng.normal(size=4)
        J = O.face_flux(Xa, Xb, Ma, Mb, g)
        M = dyadic_matrix(0.5 * (Xa + Xb), 0.5 * (Ma + Mb))
        assert np.allclose(J, M @ g, rtol=1e-13, atol=1e-14)
    # constant mobility D/RT: M = (D/RT)(diag X - X X^T)
    X = np.array([0.1, 0.2, 0.2, 0.3])
    m = 1.3777e-24
    g = rng.normal(size=4) * 1e12
    J = O.face_flux(X, X, np.full(5, m), np.full(5, m), g)
    assert np.allclose(J, m * (np.diag(X) - np.outer(X, X)) @ g, rtol=1e-13, atol=0)


@pytest.mark.parametrize("shape", [(24,), (12, 16), (8, 6, 10)])
def test_divergence_uniform_composition_fourier_mu(shape):
    h = 1e-10
    X0 = np.array([0.1, 0.2, 0.2, 0.3])
    Mob = np.full(shape + (5,), 1.3777e-24)
    X = np.broadcast_to(X0, shape + (4,)).copy()
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    m = [1, 2, 1][: len(shape)]
    phase = sum(2 * np.pi * m[d] * grids[d] / shape[d] for d in range(len(shape)))
    amp = np.array([3.0, -1.0, 2.0, 0.5]) * 1e3
    mu = np.cos(phase)[..., None] * amp
    lam = sum((2 * np.cos(2 * np.pi * m[d] / shape[d]) - 2) / h**2 for d in range(len(shape)))
    M = 1.3777e-24 * (np.diag(X0) - np.outer(X0, X0))
    expect = lam * np.cos(phase)[..., None] * (M @ amp)
    got = O.divergence_from_mu(X, Mob, mu, 1.0 / h)
    assert np.allclose(got, expect, rtol=1e-9, atol=1e-9 * np.abs(expect).max())
    # telescoping: the divergence of any field sums to zero
    rng = np.random.default_rng(6)
    Xr = X + 1e-2 * rng.normal(size=X.shape)
    d = O.divergence_from_mu(Xr, Mob * rng.uniform(0.5, 1.5, Mob.shape), rng.normal(size=mu.shape), 1.0 / h)
    assert np.abs(d.reshape(-1, 4).sum(axis=0)).max() < 1e-12 * np.abs(d).max() * d.size ** 0.5


@pytest.mark.parametrize("shape", [(20,), (12, 10), (6, 8, 10)])
def test_gradients_pm_closed_form_and_shift(shape):
    """Forward difference of cos(theta x) is -2 sin(theta/2) sin(theta (x + 1/2)) / h;
    the backward difference at c is the forward one at c - e_d (same subtraction)."""
    h = 2e-9
    dim = len(shape)
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    m = [[1, 2, 1, 3][a % 4] for a in range(4)]
    A = np.stack([np.cos(2 * np.pi * m[a] * grids[a % dim] / shape[a % dim]) for a in range(4)], axis=-1)
    gp, gm = O.gradients_pm(A, h)
    assert gp.shape == (dim,) + shape + (4,)
    for d in range(dim):
        for a in range(4):
            expect = np.zeros(shape)
            if a % dim == d:
                th = 2 * np.pi * m[a] / shape[d]
                expect = -2 * np.sin(th / 2) * np.sin(th * (grids[d] + 0.5)) / h
            assert np.allclose(gp[d][..., a], expect, atol=1e-6 / h)
        assert np.array_equal(gm[d], np.roll(gp[d], 1, axis=d))


@pytest.mark.parametrize("shape", [(16,), (8, 12), (6, 6, 8)])
def test_fluxes_pm_symmetry_and_divergence(shape):
    """fm at c is fp at c - e_d bit for bit (the face is shared), the fluxes assemble
    the fused divergence exactly (fluxes.py:393-395), and their total telescopes to 0."""
    rng = np.random.default_rng(11)
    dim = len(shape)
    h = 1e-9
    X = rng.dirichlet(np.ones(5) * 4, size=shape)[..., 1:]
    Mob = rng.uniform(0.5, 2.0, shape + (5,)) * 1e-24
    mu = rng.normal(size=shape + (4,)) * 1e3
    gp, gm = O.gradients_pm(mu, h)
    fp, fm = O.fluxes_pm(X, Mob, gp, gm)
    for d in range(dim):
        assert np.array_equal(fm[d], np.roll(fp[d], 1, axis=d))
    acc = None
    for d in range(dim):
        acc = (fp[d] - fm[d]) if acc is None else (acc + fp[d]) - fm[d]
    assert np.array_equal(acc * (1.0 / h), O.divergence_from_mu(X, Mob, mu, 1.0 / h))
    tot = (fp - fm).reshape(-1, 4).sum(axis=0)
    assert np.abs(tot).max() <= 1e-12 * np.abs(fp).max() * fp.size ** 0.5


def test_elastic_psi_closed_forms():
    from phasefield5d_amd.solver.elastic import build_khachaturyan_kernel
    from phasefield5d_amd.solver.system import build_fourier_grid_1d_along_direction, build_fourier_grid_3d
    n, dx = 64, 1e-10
    kg, kn = build_fourier_grid_1d_along_direction(n, dx, [1, 1, 0])
    B = build_khachaturyan_kernel(kg, kn, 240.0, 150.0, 120.0)
    rng = np.random.default_rng(7)
    s = rng.normal(size=n)
    # 1D: B is constant off k=0, B(0)=0 -> psi = B (s - mean s)
    assert np.allclose(O.elastic_psi(s, B), B[1] * (s - s.mean()), rtol=1e-12, atol=1e-12 * abs(B[1]))
    # 3D plane wave: psi = B(k) cos(k.x)
    shape = (8, 12, 16)
    kg, kn = build_fourier_grid_3d(*shape, dx)
    B3 = build_khachaturyan_kernel(kg, kn, 240.0, 150.0, 120.0)
    g = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    m = (1, 2, 3)
    s3 = np.cos(sum(2 * np.pi * m[d] * g[d] / shape[d] for d in range(3)))
    assert np.allclose(O.elastic_psi(s3, B3), B3[m] * s3, rtol=1e-12, atol=1e-12 * np.abs(B3).max())


def test_elastic_update_plane_wave():
    """X = X0 + a v cos(k.x): the misfit is s = a (l.v) cos(k.x), so the elastic update of
    mu = 0 is B(k) a (l.v) cos(k.x) l per component (2D in-plane and 3D grids).  A sign or
    offset slip in the misfit, a wrong component of l, or psi applied across components
    fails it."""
    from phasefield5d_amd.solver.elastic import build_khachaturyan_kernel
    from phasefield5d_amd.solver.system import build_fourier_grid_2d_in_plane, build_fourier_grid_3d
    dx = 2e-9
    X0 = np.array([0.1, 0.2, 0.3, 0.2])
    l = np.array([3.1e-3, -1.7e-3, 0.9e-3, 2.6e-3])
    v = np.array([0.7, -0.2, 0.5, 0.1])
    a = 1e-3
    for shape, m in (((16, 12), (3, 2)), ((8, 12, 16), (1, 5, 3))):
        if len(shape) == 2:
            kg, kn = build_fourier_grid_2d_in_plane(*shape, dx, [0, 0, 1])
        else:
            kg, kn = build_fourier_grid_3d(*shape, dx)
        B = build_khachaturyan_kernel(kg, kn, 240.0, 150.0, 120.0)
        g = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
        c = np.cos(sum(2 * np.pi * m[d] * g[d] / shape[d] for d in range(len(shape))))
        X = X0 + a * v * c[..., None]
        mu = np.zeros(shape + (4,))
        O.elastic_update(mu, X, X0, l, B)
        want = (B[m] * a * float(l @ v)) * c[..., None] * l
        assert abs(B[m]) > 0
        np.testing.assert_allclose(mu, want, rtol=0, atol=1e-12 * np.abs(want).max())
        # misfit alone: linear in each component with slope l_j, zero at X0
        for j in range(4):
            e = np.zeros(4)
            e[j] = 0.01
            np.testing.assert_allclose(O.misfit_scalar(np.broadcast_to(X0 + e, shape + (4,)), X0, l),
                                       0.01 * l[j], rtol=1e-12)


def test_adaptive_dt_rules():
    ctld, cfl = 10.0, 1.0
    # no change signal -> dt kept, then caps
    assert O.apply_adaptive_time(0.01, 0.0, ctld, cfl) == 0.01
    # band: lo/x <= dt <= hi/x
    assert O.apply_adaptive_time(0.01, 1.0, ctld, cfl, 1e-5, 1e-3) == 1e-3      # clipped to hi/x
    assert O.apply_adaptive_time(1e-7, 1.0, ctld, cfl, 1e-5, 1e-3) == 2e-7      # growth limited to 2x
    assert O.apply_adaptive_time(0.04, 1e-6, ctld, cfl, 1e-5, 1e-3) == 0.05     # 0.005 ctld cap
    assert O.apply_adaptive_time(0.04, 1e-6, ctld, 0.03, 1e-5, 1e-3) == 0.03    # CFL cap
    assert O.apply_adaptive_time(0.01, 1.0, ctld, cfl, 1e-3, 1e-5) == 1e-3      # swapped band


def test_clamp_and_bounds():
    X = np.array([[0.3, 0.3, 0.3, 0.3], [0.1, 0.2, 0.2, 0.3], [0.5, 0.5, 0.0, 0.0]])
    Y = O.clamp_simplex(X.copy())
    assert np.allclose(Y[0].sum(), 1.0) and np.allclose(Y[0], 0.25)
    assert np.array_equal(Y[1:], X[1:])
    assert not O.out_of_bounds(np.full((16, 4), 0.1))
    bad = np.full((16, 4), 0.1)
    bad[5, 2] = np.nan
    assert O.out_of_bounds(bad)
    bad = np.full((16, 4), 0.1)
    bad[0, 0] = -0.002
    assert O.out_of_bounds(bad)
    bad[0, 0] = -0.0005
    assert not O.out_of_bounds(bad)


@pytest.mark.parametrize("cell,hard", [
    ([0.5, 0.3, 0.2, 0.0015], True),        # sum > 1 + tol_hard (components within [0, 1])
    ([0.5, 0.3, 0.2, 0.0005], False),       # sum > 1 + tol_soft only
    ([-0.0009, -0.0009, 0.0004, 0.0], True),  # sum < -tol_hard, components above -tol_hard
    ([-0.0015, 0.1, 0.1, 0.1], True),       # component < -tol_hard
    ([-0.0002, 0.1, 0.1, 0.1], False),      # component < -tol_soft only
    ([1.0015, -0.0009, -0.0009, -0.0009], True),  # component > 1 + tol_hard, sum inside
    ([0.1, 0.1, 0.1, 0.1], False),
])
def test_out_of_bounds_branches(cell, hard):
    """system.py:12-53: each hard branch alone trips it, soft violations do not."""
    X = np.full((4, 4, 4), 0.1)
    X[2, 1] = cell
    assert O.out_of_bounds(X, 1e-5, 1e-3) is hard


def test_fluctuations_seeded_and_mean_free():
    X0 = np.array([0.1, 0.2, 0.2, 0.3])
    a = O.apply_fluctuations(X0, (50,), 1e-3, seed=11)
    b = O.apply_fluctuations(X0, (50,), 1e-3, seed=11)
    assert np.array_equal(a, b)
    d = (a - X0) / 1e-3
    # the removed per-cell mean over 5 components: the dependent noise is -sum of the 4
    dep = -d.sum(axis=1)
    assert np.all(np.abs(np.concatenate([d, dep[:, None]], axis=1)) <= 2.0 + 1e-12)


def test_fluctuations_hand_built_draw():
    """system.py:56-61 under a seed, built by hand: the legacy generator's doubles taken one
    at a time in C order, five per cell ([Fe, Mn, Ni, Co, Cu]); noise = 2 (u - 1/2), its
    mean over all five removed per cell, the last four kept.  A mean over four components,
    the first four kept, or another draw order fails it.  The package's
    apply_fluctuations (the host draw of every test problem) gives the same bits."""
    from phasefield5d_amd.solver.system import apply_fluctuations
    X0 = np.array([0.1, 0.2, 0.2, 0.3])
    shape, f, seed = (3, 5), 1e-3, 123
    rs = np.random.RandomState(seed)
    want = np.empty(shape + (4,))
    for i in range(shape[0]):
        for j in range(shape[1]):
            n = [(rs.random_sample() - 0.5) * 2.0 for _ in range(5)]
            mean = ((((n[0] + n[1]) + n[2]) + n[3]) + n[4]) / 5.0
            for k in range(4):
                want[i, j, k] = X0[k] + f * (n[k + 1] - mean)
    got = O.apply_fluctuations(X0, shape, f, seed=seed)
    assert np.array_equal(got, want)
    assert np.array_equal(apply_fluctuations(X0, shape, f, seed=seed), want)


def test_c_restatement_equals_numpy_oracle():
    from helpers import developed_state, oracle_sim, problem
    from oracle.cpu_path import CPath
    for dim, n in [(1, 160), (2, 24), (3, 8)]:
        pb = problem(dim, n)
        X = developed_state(pb, amp=0.08)
        X.reshape(-1, 4)[3] = [0.3, 0.3, 0.2, 0.15]
        ora = oracle_sim(pb)
        cp = CPath(pb.data_grid, pb.cfg.resolution, pb.tree, pb.calphad_values, pb.dx, pb.kappa, pb.X0, pb.l_vec,
                   pb.Bhat, pb.ctld, pb.dt_cfl)
        Xn, _, dtn, _, _ = ora.step(X, pb.dt0)
        Xc = X.copy()
        dtc, _ = cp.step(Xc, pb.dt0)
        assert dtc == dtn and np.array_equal(Xc, Xn)
