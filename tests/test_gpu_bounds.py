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


# one cell per case, the rest of the field at X0 = 0.1 per component (system.py:12-53)
CASES = [
    ("sum > 1 + tol_hard", [0.5, 0.3, 0.2, 0.0015], True, True),
    ("sum > 1 + tol_soft only", [0.5, 0.3, 0.2, 0.0005], False, True),
    ("sum < -tol_hard", [-0.0009, -0.0009, 0.0004, 0.0], True, True),
    ("component < -tol_hard", [-0.0015, 0.1, 0.1, 0.1], True, True),
    ("component < -tol_soft only", [-0.0002, 0.1, 0.1, 0.1], False, True),
    ("component > 1 + tol_hard", [1.0015, -0.0009, -0.0009, -0.0009], True, True),
    ("non-finite", [np.nan, 0.1, 0.1, 0.1], True, False),
    ("inside", [0.1, 0.1, 0.1, 0.1], False, False),
]


@pytest.mark.parametrize("interval", [1, 100])
@pytest.mark.parametrize("name,cell,hard,soft", CASES, ids=[c[0] for c in CASES])
def test_bounds_branches_vs_oracle(dev, name, cell, hard, soft, interval):
    """Every hard branch alone, the soft branch alone, NaN and a clean field: the drop-in
    out_of_bounds (device reduction) agrees with the oracle's, and an engine holding the
    field stops before its first step exactly when the field is hard out of bounds and
    flags the soft violations while stepping on (checks at step 0 for both intervals)."""
    from phasefield5d_amd.engine import ST_ACTIVE, ST_HARD_OOB, ST_SOFT_OOB, Engine
    from phasefield5d_amd.solver.system import out_of_bounds
    from oracle import pf5d_oracle as O
    pb = problem(2, 16, elastic=False)
    X = np.full((16, 16, 4), 0.1)
    X[7, 3] = cell
    assert O.out_of_bounds(X) is hard
    assert out_of_bounds(X) is hard
    assert out_of_bounds(torch_tensor(X)) is hard
    pb.X_init = X
    e = Engine("fp64")
    s = e.add_problem(pb, oob_interval=interval)
    e.step(3)
    st = e.state(s)
    e.close()
    if hard:
        assert st["status"] & ST_HARD_OOB and not st["status"] & ST_ACTIVE and st["steps"] == 0
    else:
        assert st["status"] & ST_ACTIVE and not st["status"] & ST_HARD_OOB and st["steps"] == 3
    assert bool(st["status"] & ST_SOFT_OOB) is soft


def torch_tensor(X):
    import torch
    return torch.from_numpy(X).cuda()
