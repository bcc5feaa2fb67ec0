"""Continuous batching: each served result equals the same request run alone."""
import numpy as np
import pytest

from helpers import problem

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _alone(pb, t_end, precision):
    from phasefield5d_amd.engine import ST_ACTIVE, Engine
    e = Engine(precision)
    s = e.add_problem(pb, t_end=t_end)
    while e.state(s)["status"] & ST_ACTIVE:
        e.step(7)
    X, st = e.field(s, "X"), e.state(s)
    e.close()
    return X, st


@pytest.mark.parametrize("precision", ["fp32", "fp64"])
def test_served_equals_alone(precision):
    from phasefield5d_amd.serve import Request, Server
    from phasefield5d_amd.engine import ST_DONE
    pbs = [problem(1, 160, seed=1), problem(2, 32, seed=2), problem(3, 16, seed=3), problem(2, 48, seed=4),
           problem(1, 96, seed=5, elastic=False)]
    reqs = []
    for i, pb in enumerate(pbs):
        steps = 15 + 9 * i
        r = Request(i, pb.cfg, None, arrival=0.05 * i, max_steps=steps)
        r.problem = pb
        r.t_end = steps * pb.dt_cfl
        reqs.append(r)
    srv = Server(precision, poll_steps=5, setup_workers=2)
    srv._setup = lambda req: req  # problems are prebuilt
    done = srv.run(reqs)
    srv.close()
    assert len(done) == len(pbs)
    for r in done:
        assert r.status & ST_DONE
        X, st = _alone(r.problem, r.t_end, precision)
        assert st["steps"] == r.steps
        assert np.array_equal(X, r.X), f"request {r.rid} differs alone vs served"


def test_admission_respects_budget():
    """With room for one request at a time, requests queue and still complete."""
    from phasefield5d_amd.serve import Request, Server, footprint_bytes
    pbs = [problem(2, 32, seed=s) for s in range(3)]
    reqs = []
    for i, pb in enumerate(pbs):
        r = Request(i, pb.cfg, None, arrival=0.0, max_steps=10)
        r.problem, r.t_end = pb, 10 * pb.dt_cfl
        reqs.append(r)
    budget = footprint_bytes(pbs[0], "fp32") + 1
    srv = Server("fp32", hbm_budget=budget, poll_steps=4)
    srv._setup = lambda req: req
    done = srv.run(reqs)
    srv.close()
    assert len(done) == 3
    spans = sorted((r.admitted, r.finished) for r in done)
    for (a0, f0), (a1, f1) in zip(spans, spans[1:]):
        assert a1 >= f0 - 1e-6  # never two resident at once
