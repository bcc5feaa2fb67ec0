"""Continuous-batching simulation server on one GPU.

Requests (dimension, size, composition, kappa, T, end time) arrive over time.
Each is set up on the host in a worker thread (table, kinetics, B(k), initial
field), admitted into a slot of one device-resident ``Engine`` as soon as its
HBM footprint fits, and stepped together with every other resident request:
each step of the engine advances all resident requests by their own adaptive
dt.  A request leaves when its device-side time passes its end time (or its
bounds check fails); its field is copied back and its slot freed, and queued
requests take the memory.  Membership changes rebuild the engine's work lists
and re-capture its step graph between polls.

Latency = completion time - arrival time (includes queueing and set-up).
A sampled fraction of requests is re-run alone on an FP64 engine and its
final composition compared with the served FP32 result ("audit").
"""
import math
import threading
import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .engine import ST_ACTIVE, ST_DONE, ST_HARD_OOB, Engine
from .problem import build_problem


@dataclass
class Request:
    rid: int
    cfg: object
    table: object
    arrival: float                  # seconds after the server starts
    t_end_ctld: float = 50.0        # end time in units of the request's CTLD
    max_steps: Optional[int] = None  # optional cap: t_end <= max_steps * dt_cfl
    audit: bool = False
    # filled by the server
    problem: object = None
    t_end: float = 0.0
    slot: int = -1
    admitted: float = math.nan
    finished: float = math.nan
    steps: int = 0
    status: int = 0
    X: Optional[np.ndarray] = None
    audit_err: Optional[float] = None
    extra: dict = field(default_factory=dict)

    @property
    def latency(self):
        return self.finished - self.arrival


def footprint_bytes(pb, precision):
    """Device bytes a request holds in an engine (fields, remainder, spectrum, kernel)."""
    w = 4 if precision == "fp32" else 8
    n = pb.ncell
    b = n * 4 * w * 2 + (n * 12 if precision == "fp32" else 0)  # X, dX (+ 24-bit remainders in FP32)
    b += n * 9 * w                                         # mu, Mob of the two-pass stencil
    if pb.Bhat is not None:
        kshape = list(pb.Bhat.shape)
        if len(kshape) > 1:                                # spectrum rows padded to 128-byte lines
            per = 128 // (2 * w)
            kshape[-1] = -(-kshape[-1] // per) * per
        nh = int(np.prod(kshape))
        b += n * 2 * w + nh * 3 * w                        # psi, s, spectrum (complex), Bhat
    return b


class Server:
    def __init__(self, precision="fp32", hbm_budget=None, poll_steps=20, setup_workers=4, autotune=False,
                 fast_init=True):
        self.engine = Engine(precision)
        self.fast_init = fast_init  # initial fields drawn on the device (Engine.add_problem)
        self.precision = precision
        if hbm_budget is None:
            import torch
            free, _ = torch.cuda.mem_get_info()
            hbm_budget = int(free * 0.7)
        self.hbm_budget = hbm_budget
        self.poll_steps = poll_steps
        self.pool = ThreadPoolExecutor(max_workers=setup_workers)
        self.autotune = autotune
        self.used_bytes = 0
        self.log = []

    def _setup(self, req):
        t_setup = time.perf_counter()
        self._setup_problem(req)
        req.extra["setup_s"] = time.perf_counter() - t_setup
        return req

    def _setup_problem(self, req):
        # audited requests take the host draw (apply_fluctuations under the request's
        # seed, the reference's seeded MT19937 sequence) so that the audit starts from
        # the reference's initial field; the others draw theirs on the device
        if self.fast_init and not req.audit:
            # the initial field is drawn on the device when the request is admitted
            pb = build_problem(req.cfg, table=req.table, X_init=np.empty((0,)))
            pb.X_init = None
            pb.extra["device_init"] = (req.cfg.seed, req.cfg.fluctuation)
        else:
            pb = build_problem(req.cfg, table=req.table)
        t_end = req.t_end_ctld * pb.ctld
        if req.max_steps is not None:
            t_end = min(t_end, req.max_steps * pb.dt_cfl)
        req.problem, req.t_end = pb, t_end
        return req

    def run(self, requests, max_wall=None, verbose=False):
        """Serve ``requests`` (arrival times relative to now); returns them completed."""
        pending = sorted(requests, key=lambda r: r.arrival)
        futures = []           # (req, future) being set up
        ready = []             # set up, waiting for memory
        running = {}           # slot -> req
        done = []
        t0 = time.perf_counter()
        i = 0
        busy = 0.0
        while len(done) < len(requests):
            now = time.perf_counter() - t0
            if max_wall is not None and now > max_wall:
                break
            while i < len(pending) and pending[i].arrival <= now:
                futures.append((pending[i], self.pool.submit(self._setup, pending[i])))
                i += 1
            still = []
            for req, fut in futures:
                if fut.done():
                    fut.result()
                    ready.append(req)
                else:
                    still.append((req, fut))
            futures = still
            # admission by HBM footprint, first come first served
            admitted = False
            while ready:
                req = ready[0]
                need = footprint_bytes(req.problem, self.precision)
                if need > self.hbm_budget:
                    raise MemoryError(f"request {req.rid} needs {need} bytes > budget {self.hbm_budget}")
                if self.used_bytes + need > self.hbm_budget:
                    break
                ready.pop(0)
                ta = time.perf_counter()
                req.slot = self.engine.add_problem(req.problem, t_end=req.t_end)
                req.admitted = time.perf_counter() - t0
                req.extra["bytes"] = need
                req.extra["admit_s"] = time.perf_counter() - ta  # table, B-hat and field on the device
                self.used_bytes += need
                running[req.slot] = req
                admitted = True
            if admitted and self.autotune:
                from . import autotune
                autotune.tune(self.engine, chunks=(8, 16), reps=3, quick=True)
            if running:
                ts = time.perf_counter()
                self.engine.step(self.poll_steps)
                states = self.engine.all_states()
                busy += time.perf_counter() - ts
                now = time.perf_counter() - t0
                for slot, st in states.items():
                    req = running.get(slot)
                    if req is None or st["status"] & ST_ACTIVE:
                        continue
                    req.steps, req.status = st["steps"], st["status"]
                    req.X = self.engine.field(slot, "X")
                    req.finished = now
                    self.engine.remove(slot)
                    self.used_bytes -= req.extra["bytes"]
                    del running[slot]
                    done.append(req)
                    if verbose:
                        print(f"  request {req.rid} {req.problem.shape} done: {req.steps} steps, "
                              f"latency {req.latency:.2f} s, status {req.status}", flush=True)
            else:
                nxt = pending[i].arrival if i < len(pending) else None
                if nxt is not None and not futures and not ready:
                    time.sleep(max(0.0, min(0.01, nxt - (time.perf_counter() - t0))))
                else:
                    time.sleep(0.001)
        self.wall = time.perf_counter() - t0
        self.busy = busy
        return done

    def close(self):
        self.pool.shutdown(wait=True)
        self.engine.close()


def audit(req, precision="fp64"):
    """Re-run a served request alone on an FP64 engine; max |X_served - X_fp64|."""
    eng = Engine(precision)
    slot = eng.add_problem(req.problem, t_end=req.t_end)
    while True:
        eng.step(200)
        st = eng.state(slot)
        if not st["status"] & ST_ACTIVE:
            break
    X = eng.field(slot, "X")
    eng.close()
    req.audit_err = float(np.abs(X - req.X).max())
    req.extra["audit_steps"] = st["steps"]
    return req.audit_err


def poisson_requests(rate, duration, seed=0, max_steps=500, audit_frac=0.05, t_end_ctld=50.0):
    """Arrivals at Poisson ``rate`` for ``duration`` s.  Dimension 1/2/3 with probability
    0.3/0.5/0.2: 1D mw 20-200 x ppw 8-16; 2D sides 128-512; 3D sides 64-256 (sides drawn
    as multiples of 16); composition and kappa as in the batched mix; end time 50 CTLD,
    capped at ``max_steps`` CFL-limited steps when given.  Every (1 / audit_frac)-th
    request is audited."""
    from .workloads import _mix_config, draw_spinodal_composition, table_for
    rng = np.random.default_rng(seed)
    t, reqs, rid = 0.0, [], 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        u = rng.uniform()
        T = int(rng.choice([873, 1000]))
        X0 = draw_spinodal_composition(rng, T)
        kappa = float(10 ** rng.uniform(-16, -14))
        if u < 0.3:
            mw, ppw = int(rng.integers(20, 201)), int(rng.integers(8, 17))
            cfg = _mix_config(1, mw * ppw, X0, kappa, T, rid, ppw=ppw)
        elif u < 0.8:
            cfg = _mix_config(2, 16 * int(rng.integers(8, 33)), X0, kappa, T, rid)
        else:
            cfg = _mix_config(3, 16 * int(rng.integers(4, 17)), X0, kappa, T, rid)
        reqs.append(Request(rid, cfg, table_for(T), t, t_end_ctld=t_end_ctld, max_steps=max_steps,
                            audit=audit_frac > 0 and rid % max(1, int(round(1.0 / audit_frac))) == 0))
        rid += 1
    return reqs


def summarize(done, wall):
    lat = np.array([r.latency for r in done])

    def pct(key, q):
        v = [r.extra[key] for r in done if key in r.extra]
        return float(np.percentile(v, q)) if v else None
    cells = sum(r.problem.ncell * r.steps for r in done)
    errs = [r.audit_err for r in done if r.audit_err is not None]
    # time from arrival to admission (host set-up, waiting for a set-up worker or for memory)
    wait = [r.admitted - r.arrival for r in done]
    return {"requests": len(done), "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
            "setup_p50_s": pct("setup_s", 50), "setup_p99_s": pct("setup_s", 99),
            "admit_p50_s": pct("admit_s", 50), "admit_p99_s": pct("admit_s", 99),
            "to_admission_p50_s": float(np.percentile(wait, 50)), "to_admission_p99_s": float(np.percentile(wait, 99)),
            "mean_s": float(lat.mean()), "wall_s": wall, "cell_updates": int(cells),
            "cell_updates_per_s": cells / wall, "audited": len(errs),
            "audit_max_err": max(errs) if errs else None,
            "failed": sum(1 for r in done if r.status & ST_HARD_OOB),
            "completed": sum(1 for r in done if r.status & ST_DONE)}


def main(argv=None):
    """Serve a Poisson stream of mixed requests for a given window and report
    latency, throughput and the FP64 audit of a sample as one JSON object:

        python -m phasefield5d_amd.serve --rate 4 --seconds 600 --max-steps 300 --json out.json

    Prints a line per completed request and per audit while it runs."""
    import argparse
    import json
    ap = argparse.ArgumentParser(description="continuous-batching serving run (one GPU, FP32)")
    ap.add_argument("--rate", type=float, default=4.0, help="mean arrivals per second")
    ap.add_argument("--seconds", type=float, default=20.0, help="arrival window")
    ap.add_argument("--max-steps", type=int, default=300, help="CFL-limited steps per request at most")
    ap.add_argument("--audit-frac", type=float, default=0.05, help="fraction of requests re-run alone in FP64")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--json", default=None, help="write the summary here as well")
    a = ap.parse_args(argv)
    reqs = poisson_requests(a.rate, a.seconds, seed=a.seed, max_steps=a.max_steps, audit_frac=a.audit_frac)
    print(f"{len(reqs)} requests over {a.seconds:g} s", flush=True)
    srv = Server("fp32", poll_steps=25, setup_workers=6)
    done = srv.run(reqs, max_wall=a.seconds * 6 + 60, verbose=True)
    srv.close()
    for r in done:
        if r.audit:
            err = audit(r)
            print(f"  audit {r.rid} {r.problem.shape}: max |X - X_fp64| {err:.3e}", flush=True)
    out = summarize(done, srv.wall)
    out.update({"arrival_rate_per_s": a.rate, "arrival_window_s": a.seconds, "submitted": len(reqs),
                "max_steps_per_request": a.max_steps, "gpu_busy_s": srv.busy,
                "audited_fraction": sum(1 for r in done if r.audit) / max(1, len(done))})
    print(json.dumps(out), flush=True)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
