"""Device parity against the FP64 oracle (oracle/pf5d_oracle.py), through the C ABI.

Bars: per-function FP64 kernels without FFT are bit-exact (same operation
order, no contraction); anything involving the FFT pair is held to 1e-12
normwise relative (the FP64 per-step tolerance); FP32 results to 1e-5 absolute in
composition after N steps.
"""
import numpy as np
import pytest

from helpers import developed_state, oracle_sim, problem, rel_inf
from oracle import pf5d_oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL64 = 1e-12


@pytest.fixture(scope="module")
def dev():
    from phasefield5d_amd import _lib
    _lib.require_gpu()
    return _lib


def test_device_info(dev):
    import ctypes as C
    out = (C.c_int * 3)()
    dev.check(dev.lib().pf5d_device_info(out))
    print("CUs, LDS/block, warp:", list(out))
    assert out[0] > 0 and out[2] == 64


@pytest.mark.parametrize("dim,n", [(1, 160), (2, 32), (3, 16)])
def test_interpolator_bitwise(dev, dim, n):
    from phasefield5d_amd.solver.interpolation import interpolator_nb
    pb = problem(dim, n)
    X = developed_state(pb, amp=0.08)
    # cells past the linear region (nearest-node fallback) and a negative component
    Xf = X.reshape(-1, 4)
    Xf[3] = [0.3, 0.3, 0.2, 0.15]
    Xf[5] = [0.33, 0.31, 0.2, 0.16]
    Xf[7] = [-0.002, 0.2, 0.3, 0.2]
    ref = O.interpolator(X, pb.data_grid, pb.cfg.resolution, pb.tree, pb.calphad_values)
    out = interpolator_nb(X, pb.data_grid, pb.cfg.resolution, pb.tree, pb.calphad_values)
    assert np.array_equal(out, ref)
    # device tensors in and out, strided output view
    Xd = torch.from_numpy(X).cuda()
    P = torch.empty(X.shape[:-1] + (9,), dtype=torch.float64, device="cuda")
    interpolator_nb(Xd, pb.data_grid, pb.cfg.resolution, pb.tree, pb.calphad_values, out=P)
    assert np.array_equal(P.cpu().numpy(), ref)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_interpolator_table_with_missing_rows(dev, prec):
    """A table with rows missing: 3 % of the nodes at random, plus the nearest node and
    its neighbours around some fallback points, so the nearest present row lies one to
    three grid steps out.  build_4d_grid leaves NaN there (interpolation.py:30-34); the
    fallback is the nearest present row, as cKDTree over the rows gives it (:203-209);
    linear cells with a missing corner stay NaN as in the reference.  Also points far
    past the simplex face (sum X = 1.3), whose nearest rows are many steps away."""
    from phasefield5d_amd.solver.interpolation import build_4d_grid, build_calphad_kdtree, interpolator_nb
    from phasefield5d_amd.problem import table_rows
    pb = problem(2, 32)
    rows = table_rows(pb.cfg)
    h = pb.cfg.resolution
    rng = np.random.default_rng(5)
    fallback = np.array([[0.31, 0.30, 0.21, 0.15], [0.331, 0.309, 0.2013, 0.1604], [0.52, 0.2, 0.15, 0.11],
                         [0.5, 0.5, 0.3, 0.0], [0.9, 0.2, 0.1, 0.1], [0.24, 0.26, 0.24, 0.2201]])
    idx = np.rint(rows[:, :4] / h).astype(int)
    drop = rng.random(len(rows)) < 0.03
    for p in fallback[:3]:
        near = np.rint(p / h).astype(int)
        drop |= np.abs(idx - near).max(axis=1) <= (1 if p[0] < 0.32 else 2)
    kept = rows[~drop]
    print(f"[holes] {prec}: {drop.sum()} of {len(rows)} rows removed", flush=True)
    _, grid = build_4d_grid(kept, h)
    tree, values = build_calphad_kdtree(kept)
    # equidistant rows are resolved by the KD-tree's traversal order in the reference,
    # which nothing specifies (the device takes the first in node order): no ties here
    d2, _ = tree.query(fallback, k=2)
    assert np.all(d2[:, 1] - d2[:, 0] > 1e-9 * h), d2 / h
    X = developed_state(pb, amp=0.08)
    Xf = X.reshape(-1, 4)
    Xf[: len(fallback)] = fallback
    Xf[len(fallback):len(fallback) + 40] = rng.uniform(0.18, 0.24, size=(40, 4))  # linear, some hit holes
    dtype = np.float64 if prec == "fp64" else np.float32
    Xp = X.astype(dtype)
    out = interpolator_nb(Xp, grid, h, tree, values).astype(np.float64)
    if prec == "fp64":
        ref = O.interpolator(Xp, grid, h, tree, values)
        fb = (((Xp[..., 0] + Xp[..., 1]) + Xp[..., 2]) + Xp[..., 3]) > O.linear_threshold(h)
        assert fb.sum() >= len(fallback) and np.isnan(ref[~fb]).any()
        assert np.array_equal(out, ref, equal_nan=True)
    else:
        # the device decides in FP32 which cells fall back; those take the nearest
        # present row (searched in FP64 from the FP32 point), stored in FP32
        fb = (((Xp[..., 0] + Xp[..., 1]) + Xp[..., 2]) + Xp[..., 3]) > np.float32(O.linear_threshold(h))
        _, nn = tree.query(Xp[fb].astype(np.float64))
        assert fb.sum() >= len(fallback)
        assert np.array_equal(out[fb], values[nn].astype(np.float32).astype(np.float64))


def test_engine_step_table_with_missing_rows(dev):
    """The fused FP64 step (elasticity off: bit-exact) on a table with rows missing
    where a patch of cells past the linear region finds its nearest rows: the engine
    takes the potentials pass whose nearest-row search skips missing rows."""
    from phasefield5d_amd.engine import Engine
    from phasefield5d_amd.problem import build_problem
    pb0 = problem(2, 32, elastic=False)
    h, num = pb0.cfg.resolution, pb0.data_grid.shape[0]
    X = developed_state(pb0, amp=0.04, seed=3)
    rng = np.random.default_rng(7)
    X[10:16, 20:26] = np.array([0.352, 0.298, 0.201, 0.118]) + rng.uniform(-0.01, 0.01, size=(6, 6, 4))
    Xf = X.reshape(-1, 4)
    lin = (((Xf[:, 0] + Xf[:, 1]) + Xf[:, 2]) + Xf[:, 3]) <= O.linear_threshold(h)
    key = lambda a: ((a[..., 0] * num + a[..., 1]) * num + a[..., 2]) * num + a[..., 3]
    rk = key(np.rint(pb0.rows[:, :4] / h).astype(int))
    used = set()  # corners the linear cells read: kept, so no linear cell turns NaN
    for x in Xf[lin]:
        i = np.clip(np.floor(x / h).astype(int), 0, num - 2)
        for b in range(16):
            used.add(int(key(i + np.array([(b >> (3 - d)) & 1 for d in range(4)]))))
    _, nn = pb0.tree.query(Xf[~lin])
    gone = set(int(k) for k in rk[nn]) - used
    pb = build_problem(pb0.cfg, table=pb0.rows[~np.isin(rk, list(gone))], X_init=X)
    _, nn2 = pb.tree.query(Xf[~lin])
    moved = (np.abs(pb0.calphad_values[nn] - pb.calphad_values[nn2]).max(axis=1) > 0).sum()
    d2, _ = pb.tree.query(Xf[~lin], k=2)
    print(f"[holes] engine: {len(gone)} rows removed, {moved} of {(~lin).sum()} fallback cells move", flush=True)
    assert moved > 0 and np.all(d2[:, 1] - d2[:, 0] > 1e-9 * h)
    Xn, dX, dt_new, _, xmax = oracle_sim(pb).step(X, pb.dt0)
    assert not np.isnan(dX).any()
    eng = Engine("fp64")
    slot = eng.add_problem(pb)
    eng.step(1)
    st = eng.state(slot)
    assert np.array_equal(eng.field(slot, "dXdt"), dX)
    assert st["dt"] == dt_new and np.array_equal(eng.field(slot, "X"), Xn)
    eng.close()


@pytest.mark.parametrize("dim,n", [(1, 160), (2, 48), (3, 24)])
def test_laplacian_and_divergence_bitwise(dev, dim, n):
    from phasefield5d_amd.solver.fluxes import get_divergence_from_mu_function
    from phasefield5d_amd.solver.operators import calculate_laplacian
    pb = problem(dim, n)
    X = developed_state(pb)
    assert np.array_equal(calculate_laplacian(X, pb.dx), O.laplacian(X, pb.dx))
    P = O.interpolator(X, pb.data_grid, pb.cfg.resolution, pb.tree, pb.calphad_values)
    Pd = torch.from_numpy(P).cuda()
    Xd = torch.from_numpy(X).cuda()
    out = torch.empty_like(Xd)
    get_divergence_from_mu_function(dim)(Xd, Pd[..., 4:], Pd[..., :4], 1.0 / pb.dx, out)
    ref = O.divergence_from_mu(X, P[..., 4:], P[..., :4], 1.0 / pb.dx)
    assert np.array_equal(out.cpu().numpy(), ref)


def test_scalar_kernels(dev):
    from phasefield5d_amd.solver.operators import abs_max, clamp_simplex, update_composition
    from phasefield5d_amd.solver.system import out_of_bounds
    rng = np.random.default_rng(3)
    a = rng.normal(size=(17, 31, 4))
    a[2, 3, 1] = np.nan
    assert abs_max(a) == O.abs_max(a)
    assert abs_max(np.zeros((5, 4))) == 0.0
    X = rng.uniform(0, 0.4, size=(1000, 4))
    ref = O.clamp_simplex(X.copy())
    got = clamp_simplex(X.copy())
    assert np.array_equal(got, ref)
    dX = rng.normal(size=X.shape)
    Xu, dXu = X.copy(), dX.copy()
    update_composition(Xu, dXu, 0.37)
    assert np.array_equal(dXu, dX * 0.37) and np.array_equal(Xu, X + dX * 0.37)
    assert not out_of_bounds(np.full((16, 4), 0.1))
    bad = np.full((16, 4), 0.1)
    bad[5, 2] = np.nan
    assert out_of_bounds(bad)
    bad = np.full((16, 4), 0.1)
    bad[3, 0] = 1.5
    assert out_of_bounds(bad)


# lengths with large prime factors (101, 193; 7 x 11 x 13) run the direct-DFT passes;
# FP64 lines of 3200 (the longest 1D request) and 4002 = 2 x 3 x 23 x 29 keep their
# twiddles in global memory; x lengths 64 and 128 and a y length of 64 take the
# compile-time plans along strided axes (lines gathered into LDS)
@pytest.mark.parametrize("dim,shape", [(1, (160,)), (1, (189,)), (1, (808,)), (1, (2316,)), (1, (1001,)), (1, (3200,)),
                                       (1, (4002,)), (2, (48, 96)),
                                       (2, (64, 64)), (2, (52, 22)), (3, (16, 24, 32)), (3, (12, 12, 12)),
                                       (3, (14, 22, 26)), (3, (64, 12, 10)), (3, (12, 64, 10)), (2, (128, 20))])
def test_elastic_update(dev, dim, shape):
    from phasefield5d_amd.solver.elastic import make_elastic_updater
    from phasefield5d_amd.solver.system import (build_fourier_grid_1d_along_direction,
                                                build_fourier_grid_2d_in_plane, build_fourier_grid_3d)
    from phasefield5d_amd.solver.elastic import build_khachaturyan_kernel
    dx = 1e-10
    if dim == 1:
        kg, kn = build_fourier_grid_1d_along_direction(shape[0], dx, [1, 1, 0])
    elif dim == 2:
        kg, kn = build_fourier_grid_2d_in_plane(shape[0], shape[1], dx, [0, 1, 1])
    else:
        kg, kn = build_fourier_grid_3d(*shape, dx)
    B = build_khachaturyan_kernel(kg, kn, 240.0, 150.0, 120.0)
    rng = np.random.default_rng(0)
    lih = np.array([0.086, 0.004, 0.008, 0.029])
    Lam = np.outer(lih, lih) * 7.1e-6

    class Cfg:
        include_cubic_anisotropy = True
        initial_composition = np.array([0.1, 0.2, 0.2, 0.3])

    X = Cfg.initial_composition + 1e-2 * rng.normal(size=shape + (4,))
    upd = make_elastic_updater(Cfg(), Lam, B, X.shape, np.float64)
    mu0 = rng.normal(size=shape + (4,)) * 1e3
    mu = mu0.copy()
    upd(mu, X)
    from phasefield5d_amd.solver.elastic import rank_one_vector
    l, _ = rank_one_vector(Lam)
    ref = O.elastic_update(mu0.copy(), X, Cfg.initial_composition, l, B)
    assert rel_inf(mu - mu0, ref - mu0) < TOL64


def _split_variant(eng):
    """Index of the per-plane two-pass stencil form (flux pass without marching)."""
    return next(i for i, v in enumerate(eng.variants()) if v[0] == 0 and v[3] == 0)


def _engine_one_step(pb, precision="fp64", form="march"):
    """One step; form "march": the engine's default tiling (marching flux pass),
    "split": the per-plane flux pass."""
    from phasefield5d_amd.engine import Engine
    eng = Engine(precision)
    slot = eng.add_problem(pb)
    if form == "split":
        eng.set_tiling(_split_variant(eng), 1)
    eng.step(1)
    eng.sync()
    return eng, slot


@pytest.mark.parametrize("form", ["march", "split"])
@pytest.mark.parametrize("dim,n,elastic", [(1, 160, True), (1, 3200, True), (2, 32, True), (2, 32, False),
                                           (3, 16, True), (3, 24, False)])
def test_fused_step_fp64(dev, dim, n, elastic, form):
    pb = problem(dim, n, elastic=elastic)
    if n == 3200:
        # the longest 1D request (its FFT keeps twiddles in global memory in FP64),
        # with the 160-cell field's wavelengths: the flux stencil differences psi
        # twice, so FFT rounding relative to dX grows as (wavelength / 2 pi)^2 and a
        # field of 1000-cell modes would measure that, not the step
        pb.X_init = np.tile(developed_state(problem(1, 160, elastic=elastic)), (20, 1))
    else:
        pb.X_init = developed_state(pb)
    ora = oracle_sim(pb)
    Xn, dX, dt_new, _, xmax = ora.step(pb.X_init, pb.dt0)
    eng, slot = _engine_one_step(pb, form=form)
    st = eng.state(slot)
    dXg = eng.field(slot, "dXdt")
    if elastic:
        assert rel_inf(dXg, dX) < TOL64
        assert abs(st["dt"] - dt_new) <= TOL64 * dt_new
        assert rel_inf(eng.field(slot, "X") - pb.X_init, Xn - pb.X_init) < 1e-10
    else:
        assert np.array_equal(dXg, dX)
        assert st["dt"] == dt_new and st["last_xmax"] == xmax
        assert np.array_equal(eng.field(slot, "X"), Xn)
    assert st["steps"] == 1 and st["t"] == pytest.approx(dt_new, rel=TOL64)


def test_fused_trajectory_per_step_fp64(dev):
    """20 device steps; each compared with one oracle step from the device's own state."""
    from phasefield5d_amd.engine import Engine
    pb = problem(2, 32)
    pb.X_init = developed_state(pb, amp=0.04)
    ora = oracle_sim(pb)
    eng = Engine("fp64")
    slot = eng.add_problem(pb)
    X, dt = pb.X_init.copy(), pb.dt0
    worst = 0.0
    for _ in range(20):
        Xn, dX, dt_new, _, _ = ora.step(X, dt)
        eng.step(1)
        st = eng.state(slot)
        worst = max(worst, rel_inf(eng.field(slot, "dXdt"), dX), abs(st["dt"] - dt_new) / dt_new)
        X, dt = eng.field(slot, "X"), st["dt"]
    assert worst < TOL64


def test_fp32_engine_error(dev):
    """FP32 step vs the FP64 oracle trajectory: max |dX| composition error after 50 steps."""
    from phasefield5d_amd.engine import Engine
    pb = problem(2, 32)
    pb.X_init = developed_state(pb, amp=0.04)
    ora = oracle_sim(pb)
    X, dt = pb.X_init.copy(), pb.dt0
    for _ in range(50):
        X, _, dt, _, _ = ora.step(X, dt)
    eng = Engine("fp32")
    slot = eng.add_problem(pb)
    eng.step(50)
    Xg = eng.field(slot, "X")
    err = np.abs(Xg - X).max()
    print("fp32 max |X - X_fp64| after 50 steps:", err)
    assert err < 1e-5
    st = eng.state(slot)
    # the FP32 state carries its rounding remainder, so the FP64 mass is held
    m0 = pb.X_init.reshape(-1, 4).sum(axis=0)
    drift = np.abs(st["mass"] - m0) / m0
    print("fp32 relative mass drift after 50 steps:", drift)
    assert drift.max() < 1e-12


def test_fp32_state_remainder_roundtrip(dev):
    """The FP32 engine holds a field as the nearest float plus a 24-bit remainder
    in units of 2^-23 of the float spacing: set_field / field round-trips every
    value to within 2^-24 of that spacing (exactly where the value has at most
    47 significant bits), across magnitudes, signs, zeros, subnormal floats and
    values on either side of powers of two."""
    from phasefield5d_amd.engine import Engine
    pb = problem(2, 16, elastic=False)
    rng = np.random.default_rng(11)
    v = rng.uniform(-0.1, 1.1, size=16 * 16 * 4)
    specials = [0.0, -0.0, 1.0, 0.5, 0.25, 2.0 ** -20, -2.0 ** -20, 1e-30, -1e-30, 1e-45, 3e-39,
                np.nextafter(0.5, 0), np.nextafter(0.5, 1), np.nextafter(0.25, 0), 1.0 - 2.0 ** -30,
                0.2 + 2.0 ** -40, 0.3 - 2.0 ** -41, 0.1 + 3 * 2.0 ** -45]
    v[:len(specials)] = specials
    X = v.reshape(16, 16, 4)
    e = Engine("fp32")
    slot = e.add_problem(pb)
    e.set_field(slot, X)
    back = e.field(slot, "X")
    e.close()
    hi = X.astype(np.float32)
    q = np.spacing(np.maximum(np.abs(hi), np.float32(np.finfo(np.float32).tiny))).astype(np.float64) * 2.0 ** -23
    err = np.abs(back - X)
    assert np.all(err <= q / 2), (err / q).max()
    # values with at most 39 significant bits come back exactly
    exact = np.round(X / q) * q == X
    assert np.array_equal(back[exact], X[exact])
    assert exact.sum() >= 8  # 0, -0, 1, 1/2, 1/4, +-2^-20, 1 - 2^-30


@pytest.mark.parametrize("precision,side", [("fp64", 520), ("fp32", 650)])
def test_engine_rejects_fields_past_32bit_offsets(dev, precision, side):
    """A simulation whose (ncell, 4) fields pass 4 GiB (the stencil kernels' 32-bit
    offsets: 2^27 cells in FP64, 2^28 in FP32) is refused with ValueError before
    anything is allocated; the engine keeps working."""
    import dataclasses
    from phasefield5d_amd.engine import Engine
    base = problem(3, 16, elastic=False)
    big = dataclasses.replace(base, shape=(side, side, side), X_init=None, Bhat=None,
                              extra={"device_init": (1, 1e-3)})
    e = Engine(precision)
    with pytest.raises(ValueError, match="field too large"):
        e.add_problem(big)
    s = e.add_problem(base)
    e.step(2)
    assert e.state(s)["steps"] == 2
    e.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_retiling_midrun_keeps_trajectory(dev, precision):
    """Changing the stencil tiling / FFT batching between steps rebuilds the work lists;
    the simulation stays active and its trajectory is unchanged (bit-identical)."""
    from phasefield5d_amd.engine import ST_ACTIVE, Engine
    pb = problem(3, 16)
    pb.X_init = developed_state(pb, amp=0.03)
    a = Engine(precision)
    sa = a.add_problem(pb)
    a.step(6)
    b = Engine(precision)
    sb = b.add_problem(pb)
    b.step(3)
    b.set_tiling(max(i for i, v in enumerate(b.variants()) if v[0] >= 0), 8)
    b.set_fft(8, 128)
    b.step(3)
    st = b.state(sb)
    assert st["status"] & ST_ACTIVE and st["steps"] == 6
    assert np.array_equal(a.field(sa, "X"), b.field(sb, "X"))
    assert np.allclose(st["mass"], pb.X_init.reshape(-1, 4).sum(axis=0), rtol=1e-12)


@pytest.mark.parametrize("shape", [(40,), (12, 10), (8, 6, 10)])
def test_gradients_and_fluxes_bitwise(dev, shape):
    """calculate_gradients_pm and get_flux_function equal the oracle bit for bit (FP64),
    for NumPy arrays and for the strided mu / Mob views of device data."""
    from phasefield5d_amd.solver.fluxes import get_flux_function
    from phasefield5d_amd.solver.operators import calculate_gradients_pm
    rng = np.random.default_rng(21)
    dim, h = len(shape), 2e-9
    X = rng.dirichlet(np.ones(5) * 4, size=shape)[..., 1:]
    P = np.concatenate([rng.normal(size=shape + (4,)) * 1e3, rng.uniform(0.5, 2.0, shape + (5,)) * 1e-24], axis=-1)
    mu, mob = P[..., :4], P[..., 4:]
    ogp, ogm = O.gradients_pm(mu, h)
    ofp, ofm = O.fluxes_pm(X, mob, ogp, ogm)
    gp, gm = calculate_gradients_pm(mu, h)
    assert np.array_equal(gp, ogp) and np.array_equal(gm, ogm)
    fp, fm = np.empty_like(gp), np.empty_like(gm)
    get_flux_function(dim)(X, mob, gp, gm, fp, fm)
    assert np.array_equal(fp, ofp) and np.array_equal(fm, ofm)
    # device tensors with row-strided views (stride 9)
    Pd, Xd = torch.from_numpy(P).cuda(), torch.from_numpy(X).cuda()
    gpd, gmd = calculate_gradients_pm(Pd[..., :4], h)
    fpd, fmd = torch.empty_like(gpd), torch.empty_like(gmd)
    get_flux_function(dim)(Xd, Pd[..., 4:], gpd, gmd, fpd, fmd)
    assert np.array_equal(fpd.cpu().numpy(), ofp) and np.array_equal(fmd.cpu().numpy(), ofm)


def test_save_step_diagnostics_vs_oracle(dev):
    """diagnostics.pre_update / post_update against NumPy sums of the oracle's fields
    (examples/simulate.py:299-346), FP64 sums in a different order: 1e-12 relative."""
    from phasefield5d_amd import diagnostics
    rng = np.random.default_rng(22)
    shape, h = (12, 10, 8), 1e-9
    kappa = np.array([1.0, 2.0, 0.5, 1.5]) * 5e-16
    X = rng.dirichlet(np.ones(5) * 4, size=shape)[..., 1:]
    P = np.concatenate([rng.normal(size=shape + (4,)) * 1e3, rng.uniform(0.5, 2.0, shape + (5,)) * 1e-24], axis=-1)
    dX = rng.normal(size=shape + (4,)) * 1e-3
    mu, mob = P[..., :4], P[..., 4:]
    gp, gm = O.gradients_pm(mu, h)
    fp, fm = O.fluxes_pm(X, mob, gp, gm)
    Pd = torch.from_numpy(P).cuda()
    pre = diagnostics.pre_update(torch.from_numpy(X).cuda(), Pd[..., :4], Pd[..., 4:], torch.from_numpy(dX).cuda(), h)
    axes = (0, 1, 2)
    assert np.allclose(pre["net"], dX.sum(axis=axes), rtol=1e-12, atol=1e-12 * np.abs(dX).sum())
    assert np.array_equal(pre["max_per_comp"], np.abs(dX).max(axis=axes))
    tot = (fp - fm).sum(axis=(0, 1, 2, 3))
    assert np.abs(pre["total_flux"] - tot).max() <= 1e-12 * np.abs(fp).sum()
    post = diagnostics.post_update(torch.from_numpy(X).cuda(), h, kappa)
    assert np.allclose(post["mass"], X.sum(axis=axes), rtol=1e-13)
    assert np.array_equal(post["comp_min"], X.min(axis=axes)) and np.array_equal(post["comp_max"], X.max(axis=axes))
    assert np.allclose(post["comp_var"], X.var(axis=axes), rtol=1e-10)
    assert post["comp_fe_min"] == 1.0 - X.sum(axis=-1).max()
    fg = -0.5 * h ** 3 * float(np.sum(X * (O.laplacian(X, h) * kappa)))
    assert abs(post["f_grad"] - fg) <= 1e-10 * abs(fg)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_stencil_forms_bitwise(dev, precision):
    """Every stencil tiling and form gives the same bits: rows
    longer than a tile (1D 600, 2D 96 = 64 + 32), ragged 3D tiles (20^3), with
    and without elasticity, in one batch."""
    from phasefield5d_amd.engine import Engine
    pbs = [problem(1, 600, seed=5), problem(2, 96, seed=6), problem(3, 20, seed=7), problem(2, 40, seed=8,
                                                                                           elastic=False)]
    for p in pbs:
        p.X_init = developed_state(p, amp=0.03)
    out = {}
    probe = Engine(precision)
    nvar = len(probe.variants())
    probe.close()
    for v in range(nvar):
        e = Engine(precision)
        slots = [e.add_problem(p) for p in pbs]
        e.set_tiling(v, 4)
        e.step(3)
        out[v] = [(e.field(s, "X"), e.field(s, "dXdt"), e.state(s)["dt"]) for s in slots]
        e.close()
    for v in range(1, nvar):
        for (Xa, da, ta), (Xb, db, tb) in zip(out[0], out[v]):
            assert np.array_equal(Xa, Xb) and np.array_equal(da, db) and ta == tb, f"variant {v}"


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fft_plane_passes_bitwise(dev, precision):
    """The plane passes (z and y transforms of one x-plane in one work item, the
    plane split into 2, 4 or 8 parts) give the same bits as the five line passes;
    where a part count needs more LDS than a block has, the line passes run."""
    from phasefield5d_amd.engine import Engine
    pb = problem(3, 128, seed=9)
    pb.X_init = developed_state(pb, amp=0.03)
    out, used = {}, {}
    # (parts, wave mask): the Stockham plane passes, then the wave-resident passes
    # (1 forward planes, 2 x lines, 4 inverse planes), alone and mixed with the others
    configs = [(0, 0), (2, 0), (4, 0), (8, 0), (0, 1), (0, 2), (0, 4), (0, 7), (4, 2), (4, 5)]
    for parts, wave in configs:
        print(f"[fft] {precision} parts {parts} wave {wave}", flush=True)
        e = Engine(precision)
        s = e.add_problem(pb)
        e.set_fft(16, 128, parts, wave=wave)
        e.step(2)
        out[parts, wave] = (e.field(s, "X"), e.field(s, "psi"), e.state(s)["dt"])
        info = e.launch_info()
        used[parts, wave] = (info["fft_parts"], info["fft_wave"])
        e.close()
    assert used[0, 0] == (0, 0) and used[4, 0] == (4, 0) and used[8, 0] == (8, 0)
    assert used[2, 0] == ((2, 0) if precision == "fp32" else (0, 0))
    assert used[0, 1] == (2, 1) and used[0, 2] == (0, 2) and used[0, 4] == (0, 4) and used[0, 7] == (2, 7)
    assert used[4, 2] == (4, 2) and used[4, 5] == (2, 5)
    for cfg in configs[1:]:
        for a, b in zip(out[0, 0], out[cfg]):
            assert np.array_equal(a, b), f"parts, wave {cfg}"


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_fft_plane_passes_mixed_lengths_bitwise(dev, precision):
    """One engine holding 64^3, 96^3 (three-pass plan) and 128^3 fields and a 2D
    field: the plane passes, one launch per plane length, give the same bits as
    the line passes for every simulation."""
    from phasefield5d_amd.engine import Engine
    pbs = [problem(3, 64, seed=11), problem(3, 96, seed=12), problem(3, 128, seed=13), problem(2, 64, seed=14)]
    for p in pbs:
        p.X_init = developed_state(p, amp=0.03)
    out = {}
    for parts, wave in ((0, 0), (4, 0), (4, 7)):
        e = Engine(precision)
        slots = [e.add_problem(p) for p in pbs]
        e.set_fft(16, 128, parts, 512, wave=wave)
        e.step(2)
        out[parts, wave] = [(e.field(s, "X"), e.field(s, "psi")) for s in slots]
        info = e.launch_info()
        assert info["fft_wave"] == wave and (info["fft_parts"] > 0) == (parts > 0)
        e.close()
    for cfg in ((4, 0), (4, 7)):
        for (Xa, pa), (Xb, pb_) in zip(out[0, 0], out[cfg]):
            assert np.array_equal(Xa, Xb) and np.array_equal(pa, pb_), cfg


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_stencil_forms_bitwise_uneven_shapes(dev, precision):
    """Every stencil tiling gives the same bits on shapes that exercise the tile
    edges: no more rows than a block has wavefronts (the periodic y wrap inside a
    tile; 4 cells is the engine's minimum per axis), a row just longer than a
    tile (a 2-cell tail), odd row lengths (a lane
    holding the row's last cell in its first slot), 37 planes (slabs of 4 or 5
    planes, so potentials items of several planes end short), a short 2D plane,
    a 1D line of odd length (elasticity off: the stencil alone)."""
    import dataclasses
    from phasefield5d_amd.engine import Engine
    base = {3: problem(3, 8, seed=21, elastic=False), 2: problem(2, 8, seed=22, elastic=False),
            1: problem(1, 8, seed=23, elastic=False)}
    pbs = []
    for shape in [(6, 4, 130), (5, 4, 33), (4, 9, 65), (37, 6, 20), (7, 200), (129,)]:
        p = dataclasses.replace(base[len(shape)], shape=tuple(shape), Bhat=None)
        p.X_init = developed_state(p, amp=0.03, seed=len(pbs) + 3)
        pbs.append(p)
    probe = Engine(precision)
    nvar = len(probe.variants())
    probe.close()
    out = {}
    for v in range(nvar):
        e = Engine(precision)
        slots = [e.add_problem(p) for p in pbs]
        e.set_tiling(v, 2)
        e.step(2)
        out[v] = [(e.field(s, "X"), e.field(s, "dXdt"), e.state(s)["dt"]) for s in slots]
        e.close()
    for v in range(1, nvar):
        for i, ((Xa, da, ta), (Xb, db, tb)) in enumerate(zip(out[0], out[v])):
            assert np.array_equal(Xa, Xb) and np.array_equal(da, db) and ta == tb, f"variant {v} shape {pbs[i].shape}"


def test_batched_equals_alone(dev):
    """Each simulation's result is bit-identical alone and inside a mixed batch."""
    from phasefield5d_amd.engine import Engine
    pbs = [problem(1, 160), problem(2, 32, seed=2), problem(3, 16, seed=3), problem(2, 48, seed=4, elastic=False)]
    for p in pbs:
        p.X_init = developed_state(p, amp=0.03)
    alone = []
    for p in pbs:
        e = Engine("fp64")
        s = e.add_problem(p)
        e.step(7)
        alone.append(e.field(s, "X"))
    mix = Engine("fp64")
    slots = [mix.add_problem(p) for p in pbs]
    mix.step(7)
    for s, ref in zip(slots, alone):
        assert np.array_equal(mix.field(s, "X"), ref)
