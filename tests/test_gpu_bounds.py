"""The device bounds check (system.py:12-53 every oob_interval steps, simulate.py:231-239).

Steps compute the FP64 bounds statistics only ahead of a check step; these tests
fix where a simulation must stop from a recorded trajectory and compare."""
import numpy as np
import pytest

from helpers import developed_state, problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from phasefield5d_amd import _lib
    _lib.require_gpu()
    return _lib


def _cmin_trajectory(pb, n):
    """min component of X_k for k = 0..n (state reads refresh the statistics)."""
    from phasefield5d_amd.engine import Engine
    e = Engine("fp64")
    s = e.add_problem(pb, oob_interval=0)
    out = [e.state(s)["comp_min"]]
    for _ in range(n):
        e.step(1)
        out.append(e.state(s)["comp_min"])
        assert out[-1] == e.field(s, "X").min()
    e.close()
    return np.array(out)


@pytest.mark.parametrize("interval", [1, 3, 4])
def test_hard_bound_stops_at_first_check_past_it(dev, interval):
    from phasefield5d_amd.engine import ST_ACTIVE, ST_HARD_OOB, Engine
    pb = problem(1, 160)
    # the lowest component (0.1, flat) develops structure through the cross terms
    # of the mobility and Hessian when the last one carries a 16-cell cosine, so the
    # smallest component value falls every step (by about 4e-7); long smooth modes
    # of every component would only decay
    X = np.broadcast_to(pb.X0, (160, 4)).copy()
    X[:, 3] += 0.02 * np.cos(2 * np.pi * np.arange(160) / 16)
    pb.X_init = X
    n = 12
    cmin = _cmin_trajectory(pb, n)
    assert np.all(np.diff(cmin) < 0), cmin
    # a threshold between the minima after steps 4 and 5; the expected stop is the
    # first check step whose field is past it
    k = 5
    c_star = 0.5 * (cmin[k - 1] + cmin[k])
    expect = next((j for j in range(0, n + 1) if j % interval == 0 and cmin[j] < c_star), None)
    assert expect is not None
    e = Engine("fp64")
    # hard when min component < c_star (tol_hard = -c_star); soft never
    s = e.add_problem(pb, oob_interval=interval, tol_soft=10.0, tol_hard=-c_star)
    e.step(n)
    st = e.state(s)
    assert st["status"] & ST_HARD_OOB and not st["status"] & ST_ACTIVE
    assert st["steps"] == expect, (st["steps"], expect, interval, cmin.tolist())
    e.close()


def test_statistics_current_after_unchecked_steps(dev):
    """After steps that skip the statistics, a state read reports those of the current field."""
    from phasefield5d_amd.engine import Engine
    pb = problem(2, 32)
    pb.X_init = developed_state(pb, amp=0.04)
    e = Engine("fp64")
    s = e.add_problem(pb)  # checks every 100 steps: steps 1..7 skip the statistics
    e.step(7)
    st = e.state(s)
    X = e.field(s, "X")
    mass = X.reshape(-1, 4).sum(axis=0)
    assert np.allclose(st["mass"], mass, rtol=1e-13, atol=0)
    assert st["comp_min"] == X.min() and st["comp_max"] == X.max()
    tot = X.sum(axis=-1)
    assert st["sum_min"] == pytest.approx(tot.min(), rel=1e-14) and st["sum_max"] == pytest.approx(tot.max(), rel=1e-14)
    e.close()
