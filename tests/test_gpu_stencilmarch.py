"""The fused marching stencil form (k_stencilmarch with the potentials pass on
the tiles' boundary rows and planes) against the two-pass form it replaces.

Every cell's mu and Mob are computed once, by one of the two kernels, with the
same arithmetic, so the fused form must give the two-pass form's bits: X, dX/dt
and dt after several FP32 steps, on the shapes it takes (3D, rows of 128
cells, axis 1 in tiles of 8 rows, axis 0 in runs of the x chunk) -- one tile
and one run (the halo rows and planes wrap into the tile), runs of 3, 4, 8
and 16 planes, developed fields (cells outside the item's node box and past
the linear region), elasticity on and off, several simulations in one engine,
cache-sized groups, and the 3D 128^3 workload itself.  Shapes it does not take
run the default tiling."""
import dataclasses

import numpy as np
import pytest

from helpers import developed_state, problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from phasefield5d_amd import _lib
    _lib.require_gpu()


def shaped(shape, seed, elastic=True, amp=0.04):
    """A 3D problem of the workload physics on an arbitrary shape (B-hat rebuilt for it)."""
    from phasefield5d_amd.solver.elastic import build_khachaturyan_kernel
    from phasefield5d_amd.solver.system import build_fourier_grid_3d
    base = problem(3, 8, seed=seed, elastic=elastic)
    Bhat = None
    if elastic:
        kg, kn = build_fourier_grid_3d(shape[0], shape[1], shape[2], base.dx)
        Bhat = build_khachaturyan_kernel(kg, kn, *base.cij)
    p = dataclasses.replace(base, shape=tuple(shape), Bhat=Bhat)
    p.X_init = developed_state(p, amp=amp, seed=seed)
    return p


def fused_index(eng):
    return next(i for i, v in enumerate(eng.variants()) if v[0] == -1)


def default_index():
    from phasefield5d_amd.engine import Engine
    e = Engine("fp32")
    v = e.launch_info()["variant"]
    e.close()
    return v


def run(pbs, variant, cx, steps, expect=None):
    from phasefield5d_amd.engine import Engine
    e = Engine("fp32")
    slots = [e.add_problem(p) for p in pbs]
    e.set_tiling(variant, cx)
    e.step(steps)
    used = e.launch_info()["variant"]
    if expect is not None:
        assert used == expect, f"engine ran variant {used}, not {expect}"
    out = [(e.field(s, "X"), e.field(s, "dXdt"), e.state(s)["dt"]) for s in slots]
    e.close()
    return out


def assert_same(a, b, what):
    for i, ((Xa, da, ta), (Xb, db, tb)) in enumerate(zip(a, b)):
        assert np.array_equal(da, db), f"{what}: dX/dt differs (simulation {i})"
        assert np.array_equal(Xa, Xb), f"{what}: X differs (simulation {i})"
        assert ta == tb, f"{what}: dt differs (simulation {i})"


@pytest.mark.parametrize("shape,cx,elastic", [
    ((16, 16, 128), 8, True),    # two runs, two tiles
    ((8, 8, 128), 8, True),      # one run, one tile: halo planes and rows wrap into it
    ((12, 16, 128), 4, False),   # runs of 4 planes, elasticity off
    ((9, 24, 128), 3, True),     # runs of 3 planes (one inner plane), three tiles
    ((32, 8, 128), 16, True),    # runs of 16 planes
])
def test_fused_march_equals_two_pass(dev, shape, cx, elastic):
    pb = shaped(shape, seed=sum(shape) + cx, elastic=elastic)
    two = run([pb], default_index(), cx, 4)
    from phasefield5d_amd.engine import Engine
    probe = Engine("fp32")
    fv = fused_index(probe)
    probe.close()
    print(f"[stencilmarch] shape {shape} cx {cx} elastic {elastic}", flush=True)
    fused = run([pb], fv, cx, 4, expect=fv)
    assert_same(two, fused, f"{shape} cx {cx}")


def test_fused_march_developed_field_and_batch(dev):
    """Three simulations in one engine, one of them with large modes (cells far
    from the tiles' node boxes, some past the linear region, on the nearest-row
    fallback)."""
    pbs = [shaped((16, 16, 128), seed=3), shaped((8, 24, 128), seed=4, amp=0.06), shaped((24, 8, 128), seed=5)]
    s = [p.X_init.reshape(-1, 4).sum(axis=1) for p in pbs]
    assert s[1].max() > 0.9, "the field should reach past the linear region"
    from phasefield5d_amd.engine import Engine
    probe = Engine("fp32")
    fv = fused_index(probe)
    probe.close()
    two = run(pbs, default_index(), 8, 3)
    fused = run(pbs, fv, 8, 3, expect=fv)
    assert_same(two, fused, "batch")


def test_fused_march_not_taken_where_it_does_not_fit(dev):
    """A shape outside the fused form's (rows of 96 cells) runs the default tiling."""
    pb = shaped((8, 8, 96), seed=7)
    from phasefield5d_amd.engine import Engine
    e = Engine("fp32")
    e.add_problem(pb)
    e.set_tiling(fused_index(e), 8)
    e.step(1)
    assert e.launch_info()["variant"] == default_index()
    e.close()


def test_fused_march_128_cubed(dev):
    """The 3D 128^3 workload: 6 steps (one graph of 4 + 2), fused vs two-pass, bit for bit;
    and the same in cache-sized groups (two simulations)."""
    from phasefield5d_amd.problem import build_problem
    from phasefield5d_amd.workloads import workload_config
    pb = build_problem(workload_config("3d_128", seed=0))
    from phasefield5d_amd.engine import Engine
    probe = Engine("fp32")
    fv = fused_index(probe)
    probe.close()
    two = run([pb], default_index(), 8, 6)
    fused = run([pb], fv, 8, 6, expect=fv)
    assert_same(two, fused, "128^3")
    pb2 = dataclasses.replace(pb, X_init=developed_state(pb, amp=0.05, seed=2))
    two = run([pb, pb2], default_index(), 8, 3)
    fused = run([pb, pb2], fv, 8, 3, expect=fv)
    assert_same(two, fused, "128^3 groups")
