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


def test_served_equals_alone_workload_sizes():
    """Requests of the serving workload's sizes (1D mw 50 and 200 x ppw 16, 2D 128^2,
    256^2 and 512^2, 3D 64^3, 96^3 and 128^3, elasticity on), arriving while the
    earlier ones run: every served FP32 result equals the request run alone, bit
    for bit."""
    from phasefield5d_amd.engine import ST_DONE
    from phasefield5d_amd.problem import build_problem
    from phasefield5d_amd.serve import Request, Server
    from phasefield5d_amd.workloads import small_config
    sizes = [(1, 800, 16), (2, 128, 16), (3, 64, 16), (2, 256, 16), (1, 3200, 16), (3, 96, 16), (2, 512, 16),
             (3, 128, 16)]
    reqs = []
    for i, (dim, n, ppw) in enumerate(sizes):
        pb = build_problem(small_config(dim, n, seed=40 + i, ppw=ppw))
        steps = 120 + 20 * i
        r = Request(i, pb.cfg, None, arrival=0.001 * i, max_steps=steps)
        r.problem = pb
        r.t_end = steps * pb.dt_cfl
        reqs.append(r)
        print(f"[serving] request {i}: {dim}D {pb.shape}, {steps} steps", flush=True)
    srv = Server("fp32", poll_steps=5, setup_workers=2)
    srv._setup = lambda req: req  # problems are prebuilt
    done = srv.run(reqs)
    srv.close()
    assert len(done) == len(sizes)
    overlapped = sum(1 for a in done for b in done if a is not b and a.admitted < b.finished and b.admitted < a.finished)
    assert overlapped > 0, "no two requests were resident together"
    for r in done:
        assert r.status & ST_DONE
        X, st = _alone(r.problem, r.t_end, "fp32")
        assert st["steps"] == r.steps
        assert np.array_equal(X, r.X), f"request {r.rid} {r.problem.shape} differs alone vs served"


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


def _splitmix_uniform(seed, i):
    """The published SplitMix64 step of counter i (Steele, Lea, Flood 2014), 53 bits."""
    m = (1 << 64) - 1
    z = (seed + (i + 1) * 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    z ^= z >> 31
    return (z >> 11) * (1.0 / 9007199254740992.0)


def test_device_drawn_initial_field():
    """A field drawn on the device: bit-equal to a restatement of the generator in
    FP64; the FP32 engine holds the same field as value + 24-bit remainder; the engine's
    FP64 mass equals the field's; a served request with a device-drawn field equals
    the same request run alone."""
    import dataclasses
    from phasefield5d_amd.engine import Engine
    base = problem(2, 16, seed=3)
    pb = dataclasses.replace(base, X_init=None, extra={"device_init": (12345, 1e-3)})
    e64, e32 = Engine("fp64"), Engine("fp32")
    s64, s32 = e64.add_problem(pb), e32.add_problem(pb)
    X64, X32 = e64.field(s64, "X"), e32.field(s32, "X")
    mass64 = e64.state(s64)["mass"]
    e64.close()
    e32.close()
    ref = np.empty((16 * 16, 4))
    for c in range(16 * 16):
        n = [(_splitmix_uniform(12345, 5 * c + k) - 0.5) * 2.0 for k in range(5)]
        m = ((((n[0] + n[1]) + n[2]) + n[3]) + n[4]) / 5.0
        ref[c] = [pb.X0[a] + 1e-3 * (n[a + 1] - m) for a in range(4)]
    assert np.array_equal(X64.reshape(-1, 4), ref)
    # the FP32 engine's value + 24-bit remainder holds the field to half a
    # quantum, 2^-24 of the float spacing at the value
    hi = X64.astype(np.float32)
    assert np.all(np.abs(X32 - X64) <= np.spacing(hi).astype(np.float64) * 2.0**-24)
    assert np.allclose(mass64, X64.reshape(-1, 4).sum(axis=0), rtol=1e-14, atol=0)
    # served with the field drawn at admission == alone
    from phasefield5d_amd.serve import Request, Server
    r = Request(0, pb.cfg, None, arrival=0.0, max_steps=20)
    r.problem, r.t_end = pb, 20 * pb.dt_cfl
    srv = Server("fp32", poll_steps=5, setup_workers=1)
    srv._setup = lambda req: req
    done = srv.run([r])
    srv.close()
    X, st = _alone(pb, r.t_end, "fp32")
    assert np.array_equal(done[0].X, X)
