"""Per-stage timing of whole steps: the timed steps advance a simulation exactly as
the captured step graph does, and every stage reports a duration."""
import numpy as np
import pytest

from helpers import problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from phasefield5d_amd import _lib
    _lib.require_gpu()
    return _lib


def _run(pb, precision, advance):
    from phasefield5d_amd.engine import Engine
    e = Engine(precision)
    s = e.add_problem(pb)
    out = advance(e)
    X, st = e.field(s, "X"), e.state(s)
    e.close()
    return X, st, out


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
@pytest.mark.parametrize("separate_dt", [False, True])
def test_stage_timing_steps_like_graph(dev, precision, separate_dt):
    pb = problem(3, 24, seed=7)
    n = 9
    X_ref, st_ref, _ = _run(pb, precision, lambda e: e.step(n))
    X, st, t = _run(pb, precision, lambda e: e.time_step_stages(n, separate_dt))
    assert st["steps"] == st_ref["steps"] == n
    assert st["dt"] == st_ref["dt"] and st["t"] == st_ref["t"]
    assert np.array_equal(X, X_ref)
    for k in ("fft", "k1_potentials", "k1_flux", "k2", "step"):
        assert t[k] > 0, (k, t)
    assert (t["dt"] > 0) == separate_dt, t
    parts = t["fft"] + t["k1_potentials"] + t["k1_flux"] + t["dt"] + t["k2"]
    assert abs(parts - t["step"]) <= 1e-3 + 1e-3 * t["step"], t


def test_stage_timing_rejects_groups(dev):
    from phasefield5d_amd.engine import Engine
    from phasefield5d_amd.problem import build_problem
    from phasefield5d_amd.workloads import workload_config
    e = Engine("fp32")
    for s in (0, 1):
        e.add_problem(build_problem(workload_config("3d_128", seed=s)))
    assert e.num_groups() == 2
    with pytest.raises(ValueError, match="more than one group"):
        e.time_step_stages(2)
    e.close()


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_kernel_timing_leaves_trajectory_unchanged(dev, precision):
    """Timing single stages between steps (and right after add, before any step) must not
    change what the steps compute: the timed stencil's max|dX/dt| and a timed dt
    controller's reduction would otherwise reach the next step's dt."""
    pb = problem(3, 24, seed=9)
    n = 6
    X_ref, st_ref, _ = _run(pb, precision, lambda e: e.step(n))

    def timed(e):
        for which in ("k1", "k1_potentials", "k1_flux", "dt", "fft", "k2prime", "step_like"):
            e.time_kernel(which, reps=3)
        e.step(3)
        for which in ("k1", "dt", "k1_flux", "step_like"):
            e.time_kernel(which, reps=2)
        e.step(n - 3)

    X, st, _ = _run(pb, precision, timed)
    assert st["steps"] == n and st["dt"] == st_ref["dt"] and st["t"] == st_ref["t"]
    assert np.array_equal(st["mass"], st_ref["mass"])
    assert np.array_equal(X, X_ref)


def test_prepare_captures_without_stepping(dev):
    """prepare(n) captures the graphs step(n) replays and runs none of them: the state is
    untouched, and the steps after it are those of an engine that never prepared."""
    pb = problem(3, 24, seed=4)
    n = 20
    X_ref, st_ref, _ = _run(pb, "fp32", lambda e: e.step(n))

    def prepared(e):
        X0 = e.field(0, "X")
        e.prepare(n)
        e.sync()
        st0 = e.state(0)
        assert st0["steps"] == 0 and np.array_equal(e.field(0, "X"), X0)
        e.step(n)

    X, st, _ = _run(pb, "fp32", prepared)
    assert st["steps"] == n and st["dt"] == st_ref["dt"] and np.array_equal(X, X_ref)
