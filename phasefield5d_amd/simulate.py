"""simulate-spinodal on MI355X: the examples/simulate.py driver with the device step.

Same flags as phasefield5d/solver/config.py (plus the additions listed there).
``--mode plain`` runs the reference loop body with one device call per reference
function (examples/simulate.py:230-320); ``--mode fused`` keeps the state on the
device and steps it with the fused engine (``--precision fp64|fp32``), reading
state back only at save points.

Outputs follow the reference's run directory (solver/io.py): data/ npz states
with keys timestep, time, dt, current_composition; the trace table
timeseries_info.csv (+ its JSON summary) with the save-step diagnostics of
examples/simulate.py:296-346, computed on the device; metadata.json.  Figures
and VTK export are not produced.  In fused mode the engine runs, between save
points, as many steps as the dt cap guarantees cannot reach the next save
time, then single steps, keeping each one's starting state for the fluxes.
"""
import time as wall

import numpy as np

from .problem import build_problem
from .solver.config import parse_args_to_config
from .solver.io import (SaveSchedule, finalize_traces, init_traces, make_problem_path, save_current_state,  # noqa: F401
                        update_traces)


def _record(traces, timestep, t, dt, dt_cfl, xmax, diag):
    update_traces(traces, timestep, t, dt, dt_cfl, diag["net"], xmax, diag["max_per_comp"], diag["mass"],
                  diag["total_flux"], diag["comp_min"], diag["comp_max"], diag["comp_fe_min"], diag["comp_var"],
                  diag["f_grad"])


def run(argv=None):
    cfg = parse_args_to_config(argv)
    import torch
    from . import _lib, diagnostics
    from .plain import PlainStepper
    _lib.require_gpu()
    t0 = wall.perf_counter()
    pb = build_problem(cfg, verbose=True)
    print(f"setup: shape {pb.shape}, dx {pb.dx:.3e} m, ctld {pb.ctld:.4e} s, dt0 {pb.dt0:.3e}, "
          f"dt_cfl {pb.dt_cfl:.3e} ({wall.perf_counter() - t0:.1f} s)", flush=True)
    run_dir = make_problem_path(cfg, pb) if cfg.save_outputs else None

    def save_state(step, dt, t, X, postfix=""):
        if run_dir is not None:
            save_current_state(run_dir, step, dt, t, X, postfix)

    save_state(0, 0.0, 0.0, pb.X_init, "_initial")
    sched = SaveSchedule(pb.ctld, cfg.total_timesteps, cfg.early_stage_frames, cfg.late_burst_duration_ctld,
                         cfg.late_burst_frames, cfg.late_tail_frames)
    traces = init_traces(4)
    start = wall.perf_counter()
    steps_done = 0
    last = 0
    if cfg.mode == "plain":
        ps = PlainStepper(pb, np.float64)
        for step in range(cfg.total_timesteps):
            last = step
            if not ps.step(save=sched):
                print(f"hard out-of-bounds before step {step}; stopping", flush=True)
                save_state(step, ps.dt, ps.t, ps.X.cpu().numpy(), "_out_of_bound")
                break
            steps_done += 1
            if ps.last_diag is not None:
                _record(traces, step, ps.t, ps.dt, pb.dt_cfl, ps.xmax, ps.last_diag)
                save_state(step, ps.dt, ps.t, ps.X.cpu().numpy())
        X_final, t_final, dt_final = ps.X.cpu().numpy(), ps.t, ps.dt
    else:
        from .engine import ST_ACTIVE, ST_HARD_OOB, Engine
        eng = Engine(cfg.precision)
        slot = eng.add_problem(pb)
        try:
            from . import autotune
            autotune.tune(eng, reps=3)
        except Exception as e:  # tuning is an optimisation only
            print(f"autotune skipped: {e}")
        helper = PlainStepper(pb, np.float64)  # mu / Mob of a state, for the save-step fluxes
        dt_cap = min(0.005 * pb.ctld, pb.dt_cfl)  # no step advances time by more (system.py:90-112)
        t, dt = 0.0, pb.dt0
        while steps_done < cfg.total_timesteps:
            last = steps_done
            n = sched.steps_without_save(steps_done, t, dt_cap)
            X_pre = None
            if n == 0:
                X_pre = eng.field(slot, "X")  # this step may save: keep the state it starts from
            eng.step(max(1, n))
            st = eng.state(slot)
            steps_done, t, dt = st["steps"], st["t"], st["dt"]
            if not st["status"] & ST_ACTIVE:
                print(f"simulation stopped at step {steps_done} (status {st['status']})", flush=True)
                if st["status"] & ST_HARD_OOB:
                    save_state(steps_done, dt, t, eng.field(slot, "X"), "_out_of_bound")
                break
            if X_pre is None:
                continue
            step = steps_done - 1
            sd, _ = sched(step, t)
            if not sd:
                continue
            helper.X.copy_(torch.from_numpy(X_pre))
            mu, mob = helper.potentials()
            dXdt = torch.from_numpy(eng.field(slot, "dXdt")).cuda()
            X_post = eng.field(slot, "X")
            diag = diagnostics.pre_update(helper.X, mu, mob, dXdt, pb.dx)
            diag.update(diagnostics.post_update(torch.from_numpy(X_post).cuda(), pb.dx, helper.kappa))
            _record(traces, step, t, dt, pb.dt_cfl, st["last_xmax"], diag)
            save_state(step, dt, t, X_post)
        X_final, t_final, dt_final = eng.field(slot, "X"), t, dt
    elapsed = wall.perf_counter() - start
    save_state(last, dt_final, t_final, X_final)
    if run_dir is not None and traces["timestep"]:
        finalize_traces(run_dir, traces)
    rate = np.prod(pb.shape) * steps_done / max(elapsed, 1e-12)
    print(f"Simulation complete: {steps_done} steps, t = {t_final:.4e} s, {elapsed:.3f} s "
          f"({rate:.3e} cell-updates/s). Files in {run_dir if run_dir else '(not saved)'}", flush=True)
    return {"steps": steps_done, "time": t_final, "dt": dt_final, "X": X_final, "run_dir": run_dir,
            "traces": traces, "elapsed": elapsed}


def main():
    run()


if __name__ == "__main__":
    main()
