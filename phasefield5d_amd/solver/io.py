"""Run directory, state files, save schedule and trace table (phasefield5d/solver/io.py).

Produces the reference's run layout: ``<root>/<label>_at_<T>_<dim>dim/<model>/
cells..._dx.._fluc.._kappa.._steps.._dt..s_<timestamp>/`` with ``data/`` (npz states
``step_<9 digits><postfix>.npz`` holding timestep, time, dt, current_composition),
``snapshots/`` (figures are not drawn here), ``metadata.json``, and at the end
``timeseries_info.csv`` plus its ``timeseries_info.json`` summary.
"""
import datetime
import json
import os

import numpy as np

from ..utils.labels import get_composition_label


def get_model_tag(atomic_radius_tag, include_cubic_anisotropy, direction, system_dim):
    """io.py:24-35: 'chemical', or 'elastic_cubic' with the 1D direction / 2D normal."""
    if not include_cubic_anisotropy:
        return "chemical"
    model = "elastic_cubic"
    s = "".join(map(str, direction))
    if system_dim == 1:
        model += f"_direction{s}"
    elif system_dim == 2:
        model += f"_normal{s}"
    return model


class SaveSchedule:
    """Time-based save schedule of io.py:42-109: early / burst / tail spacing in CTLD,
    plus every-N-step saves; ``__call__(step, t_after)`` -> (save_data, save_snapshot)."""

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

    def steps_without_save(self, step, t, dt_cap):
        """How many steps starting at index ``step`` (time ``t`` before it) are certain not to
        save, when no step advances time by more than ``dt_cap``; the schedule's state only
        changes on saving steps, so those steps need not be shown to it."""
        if step == 0:
            return 0
        n = self.total - 1 - step  # the last step always saves
        if self.every_data:
            r = step % self.every_data
            n = min(n, 0 if r == 0 else self.every_data - r)
        if dt_cap > 0:
            n = min(n, int(max(0.0, self.next_data - t) // dt_cap) - 1)
        return max(0, n)


def should_save(step, ctld, time, total_steps, *, early_frames=100, burst_duration_ctld=100.0, burst_frames=800,
                tail_dt_ctld=10.0, snapshot_factor=10, save_every_n_steps_data=50_000,
                save_every_n_steps_snap=100_000):
    """io.py:42-109: (save_data, save_snapshot) for loop index ``step`` at ``time``.

    Stateful like the reference: step 0 (or the first call) starts a new schedule,
    kept in the function attribute ``should_save._state`` (a dict with the phase
    ends, the data / snapshot spacings of each phase and the next save times), and
    later calls advance it.  The spacing rules are those of ``SaveSchedule``.
    """
    st = getattr(should_save, "_state", None)
    if st is None or step == 0:
        sched = SaveSchedule(ctld, total_steps, early_frames, burst_duration_ctld, burst_frames, tail_dt_ctld,
                             snapshot_factor, save_every_n_steps_data, save_every_n_steps_snap)
        should_save._state = vars(sched)
        should_save._sched = sched
        return True, True
    sched = should_save._sched
    # the reference reads its limits from the state it made at step 0; the call's
    # own total_steps and every-N values apply from this call on
    sched.total = total_steps
    sched.every_data, sched.every_snap = save_every_n_steps_data, save_every_n_steps_snap
    return sched(step, time)


def _results_root(cfg):
    return getattr(cfg, "results_root", None) or os.path.join(os.getcwd(), "results")


def make_path(cfg, hessian_max, mobility_max, number_of_cells, cell_size, system_length, ctld, wavenumber_max,
              wavelength_max, time_increment, root=None):
    """io.py:114-198 (make_path :193): create the run directory tree and metadata.json; return its path.

    Same keyword arguments as the reference; ``root`` (added) overrides the results
    root (``cfg.results_root`` if the config has one, else ./results).
    """
    root = root or _results_root(cfg)
    label = get_composition_label(cfg.initial_composition, unicode=False, elements=cfg.elements)
    sub = os.path.join(root, f"{label}_at_{cfg.temperature}_{cfg.system_dim}dim",
                       get_model_tag(cfg.atomic_radius_tag, cfg.include_cubic_anisotropy, cfg.direction,
                                     cfg.system_dim))
    stamp = datetime.datetime.now().strftime("%Y%m%d_%H%M%S")
    cells = "x".join(str(n) for n in number_of_cells)
    name = (f"cells{cells}_dx{cell_size:.0e}_fluc{cfg.fluctuation:.0e}_kappa{cfg.kappa_value:.0e}"
            f"_steps{cfg.total_timesteps}_dt{time_increment:.0e}s_{stamp}")
    run_dir = os.path.abspath(os.path.join(sub, name))
    for d in ("snapshots", "data"):
        os.makedirs(os.path.join(run_dir, d), exist_ok=True)
    meta = {
        "elements": cfg.elements, "temperature": cfg.temperature,
        "initial_composition": np.asarray(cfg.initial_composition).tolist(),
        "hessian_max": hessian_max, "mobility_max": mobility_max, "system_dim": cfg.system_dim,
        "include_cubic_anisotropy": cfg.include_cubic_anisotropy, "direction": np.asarray(cfg.direction).tolist(),
        "atomic_radius": cfg.atomic_radius_tag, "mw": cfg.mw, "ppw": cfg.ppw,
        "number_of_cells": list(number_of_cells), "cell_size": cell_size, "system_length": system_length,
        "fluctuation": cfg.fluctuation, "kappa": cfg.kappa_value, "total_timesteps": cfg.total_timesteps,
        "ctld": ctld, "wavenumber_max": wavenumber_max, "wavelength_max": wavelength_max,
        "time_steps_per_ctld": cfg.steps_per_ctld, "time_increment": time_increment, "safety": cfg.safety,
        "lower_limit": cfg.lower_limit, "upper_limit": cfg.upper_limit,
        "time_linear_end_multiplier": cfg.time_linear_end_multiplier,
        "early_stage_frames": cfg.early_stage_frames, "late_burst_duration_ctld": cfg.late_burst_duration_ctld,
        "late_burst_frames": cfg.late_burst_frames, "late_tail_frames": cfg.late_tail_frames,
        "threads": cfg.threads, "timestamp": stamp, "run_dir": run_dir,
    }
    with open(os.path.join(run_dir, "metadata.json"), "w") as f:
        json.dump(meta, f, indent=4)
    print(f"Simulation path created at:\n → {run_dir}")
    return run_dir


def make_problem_path(cfg, pb, root=None):
    """``make_path`` with the set-up values taken from a built problem (``problem.build_problem``)."""
    from .system import get_system_dimensions
    wavelength, _, system_length, _ = get_system_dimensions(pb.k_max, ppw=cfg.ppw, mw=cfg.mw, dim=cfg.system_dim)
    return make_path(cfg=cfg, hessian_max=pb.hessian_max, mobility_max=pb.mobility_max, number_of_cells=pb.shape,
                     cell_size=pb.dx, system_length=system_length, ctld=pb.ctld, wavenumber_max=pb.k_max,
                     wavelength_max=wavelength, time_increment=pb.dt0, root=root)


_FIGURE_NOTE = []


def save_snapshot_figure(*args, **kwargs):
    """io.py:223 draws matplotlib snapshot PNGs.  Figures are not part of this build:
    the call only says so, once, and the state files and traces carry the data."""
    if not _FIGURE_NOTE:
        _FIGURE_NOTE.append(True)
        print("save_snapshot_figure: snapshot figures are not drawn by phasefield5d_amd "
              "(the data/ npz states hold the fields)")


def save_current_state(path, timestep, time_increment, time, current_composition, postfix=""):
    """io.py:205-213: data/step_<timestep:09d><postfix>.npz."""
    np.savez(os.path.join(path, "data", f"step_{timestep:09d}{postfix}.npz"),
             timestep=np.int64(timestep), time=np.float64(time), dt=np.float64(time_increment),
             current_composition=np.asarray(current_composition, dtype=np.float64))


def init_traces(n_components):
    """io.py:287-305: the trace columns."""
    return {
        "timestep": [], "time": [], "dt": [], "t_cfl": [], "cfl_ratio": [], "net": [],
        "composition_change_max": [],
        **{f"composition_change_max_s{j}": [] for j in range(n_components)},
        "mass": [], "total_flux": [],
        **{f"comp_min_s{j}": [] for j in range(n_components)},
        **{f"comp_max_s{j}": [] for j in range(n_components)},
        "comp_fe_min": [],
        **{f"comp_var_s{j}": [] for j in range(n_components)},
        "f_grad": [],
    }


def update_traces(traces, timestep, time, dt, t_cfl, net, composition_change_max,
                  composition_change_max_per_solute, mass, total_flux, comp_min, comp_max, comp_fe_min, comp_var,
                  f_grad):
    """io.py:308-329: append one save step."""
    traces["timestep"].append(int(timestep))
    traces["time"].append(float(time))
    traces["dt"].append(float(dt))
    traces["t_cfl"].append(float(t_cfl))
    traces["cfl_ratio"].append(float(dt / t_cfl) if t_cfl > 0.0 else float("inf"))
    traces["net"].append(np.asarray(net).tolist())
    traces["composition_change_max"].append(float(composition_change_max))
    for j, v in enumerate(np.asarray(composition_change_max_per_solute).ravel()):
        traces[f"composition_change_max_s{j}"].append(float(v))
    traces["mass"].append(np.asarray(mass).tolist())
    traces["total_flux"].append(np.asarray(total_flux).tolist())
    for j, v in enumerate(np.asarray(comp_min).ravel()):
        traces[f"comp_min_s{j}"].append(float(v))
    for j, v in enumerate(np.asarray(comp_max).ravel()):
        traces[f"comp_max_s{j}"].append(float(v))
    traces["comp_fe_min"].append(float(comp_fe_min))
    for j, v in enumerate(np.asarray(comp_var).ravel()):
        traces[f"comp_var_s{j}"].append(float(v))
    traces["f_grad"].append(float(f_grad))


def finalize_traces(path, traces, filename="timeseries_info.csv"):
    """io.py:332-338: the CSV table and a JSON summary of its rows and columns."""
    import pandas as pd
    df = pd.DataFrame(traces)
    out = os.path.join(path, filename)
    df.to_csv(out, index=False)
    with open(os.path.join(path, "timeseries_info.json"), "w") as f:
        json.dump({"rows": len(df), "columns": list(df.columns)}, f, indent=2)
    return out
