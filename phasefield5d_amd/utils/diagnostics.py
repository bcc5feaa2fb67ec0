"""Console diagnostics the reference driver prints (phasefield5d/utils/diagnostics.py:43-74)."""
import numpy as np


def print_time_diagnositcs(time_increment, ctld, time_increment_cfl, target_steps, safety):
    """The 'Time Step Diagnostics' block: steps per CTLD asked for and obtained, the chosen
    dt beside the CFL-stable one (name spelled as in the reference)."""
    lines = [
        "",
        "--- Time Step Diagnostics ---",
        "Early-stage resolution:",
        f"  - Requested steps per CTLD:       {target_steps:_}",
        f"  - Actual steps per CTLD:          {int(ctld / time_increment):_}",
        "",
        "Time step stability limits:",
        f"  - Chosen dt:                       {time_increment:.1e} s",
        f"  - CFL-stable dt (safety={safety}): {time_increment_cfl:.1e} s",
        f"  - Steps per CTLD if using dt_cfl: {int(ctld / time_increment_cfl):_}",
        "--------------------------------",
        "",
    ]
    print("\n".join(lines))


def debugging_update(current_composition, timestep, time_increment, composition_change_max, axes_spatial):
    """What the driver prints when a bounds check fails: the step, dt, max rate, and the
    field's per-component extremes with and without non-finite values."""
    X = np.asarray(current_composition)
    print(f"Composition is out of bounds at timestep {timestep}.")
    print(f"Time increment {time_increment}s with a maxima composition change of {composition_change_max}/s.")
    print("Out of bounds before interpolation!")
    print("finite?", np.isfinite(X).all())
    print("min (w/ Nan and inf):", np.min(X, axis=axes_spatial))
    print("max (w/ Nan and inf):", np.max(X, axis=axes_spatial))
    print("min:", np.nanmin(X, axis=axes_spatial))
    print("max:", np.nanmax(X, axis=axes_spatial))
    sums = X.sum(axis=-1)
    print("sum min:", np.nanmin(sums))
    print("sum max:", np.nanmax(sums))
