"""simulate-spinodal on MI355X: the examples/simulate.py driver with the device step.

Same flags as phasefield5d/solver/config.py (plus the additions listed there).
``--mode plain`` runs the reference loop body with one device call per reference
function (examples/simulate.py:230-320); ``--mode fused`` keeps the state on the
device and steps it with the fused engine (``--precision fp64|fp32``), reading
state back only at save points.

Outputs follow the reference's run directory contents (solver/io.py): npz
states with keys timestep, dt, time, current_composition; a trace table
(CSV + JSON) with the save-step diagnostics; metadata.json.  Figures and VTK
export are not produced.
"""
import json
import os
import sys
import time as wall

import numpy as np

from .problem import build_problem
from .solver.config import parse_args_to_config


class SaveSchedule:
    """Time-based save schedule of solver/io.py:42-109 (early / burst / tail spacing in CTLD)."""

    def __init__(self, ctld, total_steps, early_frames=100, burst_duration_ctld=100.0, burst_frames=800,
                 tail_dt_ctld=10.0, snapshot_factor=10, every_data=50_000, every_snap=100_000):
        self.total = total_steps
        self.t_early_end = ctld
        self.t_burst_end = ctld * (1.0 + burst_duration_ctld)
        ef = max(1, early_frames)
        self.dt_data_early = ctld / ef
        self.dt_snap_early = ctld / max(1, ef // snapshot_factor)
        bf = max(1, burst_frames)
        burst = self.t_burst_end - self.t_early_end
        self.dt_data_burst = burst / bf
        self.dt_snap_burst = burst / max(1, bf // snapshot_factor)
        self.dt_data_tail = max(1e-30, tail_dt_ctld * ctld)
        self.dt_snap_tail = self.dt_data_tail * snapshot_factor
        self.next_data = 0.0
        self.next_snap = 0.0
        self.every_data, self.every_snap = every_data, every_snap

    def __call__(self, step, t):
        if step == 0:
            return True, True
        if step == self.total - 1:
            return True, True
        if t < self.t_early_end:
            dd, ds = self.dt_data_early, self.dt_snap_early
        elif t < self.t_burst_end:
            dd, ds = self.dt_data_burst, self.dt_snap_burst
        else:
            dd, ds = self.dt_data_tail, self.dt_snap_tail
        sd = ss = False
        if t >= self.next_data:
            while t >= self.next_data:
                self.next_data += dd
            sd = True
        if t >= self.next_snap:
            while t >= self.next_snap:
                self.next_snap += ds
            ss = True
        if self.every_data and step % self.every_data == 0:
            sd = True
        if self.every_snap and step % self.every_snap == 0:
            ss = True
        return sd, ss


def _grad_energy(Xd, h, kappa, dim):
    """f_grad = -(h^dim / 2) sum X . kappa lap X (simulate.py:336-345), Laplacian on the device."""
    from .solver.operators import calculate_laplacian
    L = calculate_laplacian(Xd, h)
    L = L * Xd.new_tensor(np.asarray(kappa).reshape(-1))
    return float(-0.5 * h**dim * (Xd * L).sum().item())


def run(argv=None):
    cfg = parse_args_to_config(argv)
    import torch
    from . import _lib
    _lib.require_gpu()
    t0 = wall.perf_counter()
    pb = build_problem(cfg, verbose=True)
    dim = len(pb.shape)
    print(f"setup: shape {pb.shape}, dx {pb.dx:.3e} m, ctld {pb.ctld:.4e} s, dt0 {pb.dt0:.3e}, "
          f"dt_cfl {pb.dt_cfl:.3e} ({wall.perf_counter() - t0:.1f} s)", flush=True)
    tag = "elastic_cubic" if cfg.include_cubic_anisotropy else "chemical"
    run_dir = os.path.join(cfg.results_root, f"{tag}_{dim}dim_{cfg.mode}_{cfg.precision}_{int(wall.time())}")
    if cfg.save_outputs:
        os.makedirs(run_dir, exist_ok=True)
        meta = {"shape": list(pb.shape), "cell_size": pb.dx, "ctld": pb.ctld, "wavenumber_max": pb.k_max,
                "time_increment": pb.dt0, "time_increment_cfl": pb.dt_cfl, "hessian_max": pb.hessian_max,
                "mobility_max": pb.mobility_max, "mode": cfg.mode, "precision": cfg.precision,
                "argv": list(argv) if argv is not None else sys.argv[1:]}
        with open(os.path.join(run_dir, "metadata.json"), "w") as f:
            json.dump(meta, f, indent=1)

    def save_state(step, dt, t, X, postfix=""):
        if cfg.save_outputs:
            np.savez(os.path.join(run_dir, f"state_{step:09d}{postfix}.npz"), timestep=step, dt=dt, time=t,
                     current_composition=X)

    save_state(0, 0.0, 0.0, pb.X_init, "_initial")
    sched = SaveSchedule(pb.ctld, cfg.total_timesteps, cfg.early_stage_frames, cfg.late_burst_duration_ctld,
                         cfg.late_burst_frames, cfg.late_tail_frames)
    traces = []
    kappa = pb.kappa.reshape(-1)
    start = wall.perf_counter()
    steps_done = 0
    if cfg.mode == "plain":
        from .plain import PlainStepper
        ps = PlainStepper(pb, np.float64)
        for step in range(cfg.total_timesteps):
            if not ps.step():
                print(f"hard out-of-bounds before step {step}; stopping", flush=True)
                save_state(step, ps.dt, ps.t, ps.X.cpu().numpy(), "_out_of_bound")
                break
            steps_done += 1
            sd, _ = sched(step, ps.t)
            if sd:
                X = ps.X
                Xh = X.cpu().numpy()
                axes = tuple(range(dim))
                traces.append({"timestep": step, "time": ps.t, "dt": ps.dt, "dt_cfl": pb.dt_cfl,
                               "max_change": ps.xmax,
                               **{f"mass_{i}": float(v) for i, v in enumerate(Xh.sum(axis=axes))},
                               **{f"min_{i}": float(v) for i, v in enumerate(Xh.min(axis=axes))},
                               **{f"max_{i}": float(v) for i, v in enumerate(Xh.max(axis=axes))},
                               **{f"var_{i}": float(v) for i, v in enumerate(Xh.var(axis=axes))},
                               "comp_fe_min": 1.0 - float(Xh.sum(axis=-1).max()),
                               "f_grad": _grad_energy(X, pb.dx, kappa, dim)})
                save_state(step, ps.dt, ps.t, Xh)
        X_final, t_final, dt_final = ps.X.cpu().numpy(), ps.t, ps.dt
    else:
        from .engine import ST_ACTIVE, Engine
        eng = Engine(cfg.precision)
        slot = eng.add_problem(pb)
        try:
            from . import autotune
            autotune.tune(eng, reps=3)
        except Exception as e:  # tuning is an optimisation only
            print(f"autotune skipped: {e}")
        chunk = 100
        while steps_done < cfg.total_timesteps:
            n = min(chunk, cfg.total_timesteps - steps_done)
            eng.step(n)
            st = eng.state(slot)
            steps_done = st["steps"]
            if not st["status"] & ST_ACTIVE:
                print(f"simulation stopped at step {steps_done} (status {st['status']})", flush=True)
                break
            sd, _ = sched(steps_done - 1, st["t"])
            if sd:
                Xh = eng.field(slot, "X")
                Xd = torch.from_numpy(Xh).cuda()
                axes = tuple(range(dim))
                traces.append({"timestep": steps_done - 1, "time": st["t"], "dt": st["dt"], "dt_cfl": pb.dt_cfl,
                               "max_change": st["last_xmax"],
                               **{f"mass_{i}": float(v) for i, v in enumerate(st["mass"])},
                               **{f"min_{i}": float(v) for i, v in enumerate(Xh.min(axis=axes))},
                               **{f"max_{i}": float(v) for i, v in enumerate(Xh.max(axis=axes))},
                               **{f"var_{i}": float(v) for i, v in enumerate(Xh.var(axis=axes))},
                               "comp_fe_min": 1.0 - float(Xh.sum(axis=-1).max()),
                               "f_grad": _grad_energy(Xd, pb.dx, kappa, dim)})
                save_state(steps_done - 1, st["dt"], st["t"], Xh)
        st = eng.state(slot)
        X_final, t_final, dt_final = eng.field(slot, "X"), st["t"], st["dt"]
    elapsed = wall.perf_counter() - start
    save_state(steps_done, dt_final, t_final, X_final)
    if cfg.save_outputs and traces:
        import pandas as pd
        df = pd.DataFrame(traces)
        df.to_csv(os.path.join(run_dir, "timeseries_info.csv"), index=False)
        df.to_json(os.path.join(run_dir, "timeseries_info.json"), orient="records")
    rate = np.prod(pb.shape) * steps_done / max(elapsed, 1e-12)
    print(f"Simulation complete: {steps_done} steps, t = {t_final:.4e} s, {elapsed:.3f} s "
          f"({rate:.3e} cell-updates/s). Files in {run_dir if cfg.save_outputs else '(not saved)'}", flush=True)
    return {"steps": steps_done, "time": t_final, "dt": dt_final, "X": X_final, "run_dir": run_dir,
            "traces": traces, "elapsed": elapsed}


def main():
    run()


if __name__ == "__main__":
    main()
