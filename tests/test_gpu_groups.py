"""Simulations whose working sets together exceed the last-level cache step group by group.

A group runs several steps while its fields stay in the cache, then the next one;
every simulation must come out bit-identical to the same simulation stepped alone."""
import numpy as np
import pytest

from helpers import developed_state, problem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from phasefield5d_amd import _lib
    _lib.require_gpu()
    return _lib


def _alone(pb, precision, steps):
    from phasefield5d_amd.engine import Engine
    e = Engine(precision)
    s = e.add_problem(pb)
    assert e.num_groups() == 1
    e.step(steps)
    X, st = e.field(s, "X"), e.state(s)
    e.close()
    return X, st


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_groups_bitwise_equal_alone(dev, precision):
    from phasefield5d_amd.engine import Engine
    from phasefield5d_amd.problem import build_problem
    from phasefield5d_amd.workloads import workload_config
    # two 128^3 fields (each one group: its working set alone is about the group
    # budget), then ten small 2D fields (one group of more than eight simulations,
    # whose dt controllers run as their own launch)
    pbs = [build_problem(workload_config("3d_128", seed=s)) for s in (0, 1)]
    for k in range(10):
        p = problem(2, 32, seed=40 + k)
        p.X_init = developed_state(p, amp=0.03, seed=k)
        pbs.append(p)
    steps = 19  # one graph of 16 steps per group, then 3 single steps
    e = Engine(precision)
    slots = [e.add_problem(p) for p in pbs]
    assert e.num_groups() == 3
    e.step(steps)
    got = [(e.field(s, "X"), e.state(s)) for s in slots]
    e.close()
    for i, p in enumerate(pbs):
        Xa, sta = _alone(p, precision, steps)
        Xb, stb = got[i]
        assert stb["steps"] == steps and sta["steps"] == steps
        assert np.array_equal(Xa, Xb), f"simulation {i} {p.shape}"
        assert sta["dt"] == stb["dt"] and sta["t"] == stb["t"]
        assert np.array_equal(sta["mass"], stb["mass"])


def test_groups_rebuilt_after_removal(dev):
    """Removing simulations between steps regroups the rest; each survivor's
    trajectory equals the same simulation stepped alone."""
    from phasefield5d_amd.engine import Engine
    from phasefield5d_amd.problem import build_problem
    from phasefield5d_amd.workloads import workload_config
    pbs = [build_problem(workload_config("3d_128", seed=s)) for s in (0, 1, 2)]
    for k in range(3):
        p = problem(2, 32, seed=60 + k)
        p.X_init = developed_state(p, amp=0.03, seed=k)
        pbs.append(p)
    e = Engine("fp32")
    slots = [e.add_problem(p) for p in pbs]
    assert e.num_groups() == 4  # one per 128^3 field, one for the three small ones
    e.step(17)
    e.remove(slots[1])  # a 128^3 group disappears
    e.remove(slots[4])  # a small one from the last group
    assert e.num_groups() == 3
    e.step(6)
    keep = [0, 2, 3, 5]
    got = {i: (e.field(slots[i], "X"), e.state(slots[i])) for i in keep}
    e.close()
    for i in keep:
        Xa, sta = _alone(pbs[i], "fp32", 23)
        Xb, stb = got[i]
        assert stb["steps"] == 23
        assert np.array_equal(Xa, Xb), f"simulation {i} {pbs[i].shape}"
        assert sta["dt"] == stb["dt"] and sta["t"] == stb["t"]
