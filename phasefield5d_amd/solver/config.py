"""Run configuration: dataclass + argparse parser.

The command line of phasefield5d/solver/config.py:74-143 -- every flag name,
default and group, the ``--x/--no_x`` boolean pairs, comma (or semicolon)
separated number lists of a fixed length, ``direction`` truncated to integers,
exactly five elements -- described here as one table (``FLAGS``) that the
parser is built from.  Added flags (all defaulting to the reference behaviour):

* ``--c11/--c12/--c44`` (GPa) override the hard-coded cubic constants;
* ``--molar_volume`` (m^3/mol) overrides the linear-mixing molar volume;
* ``--seed`` seeds the initial fluctuation draw (the reference draw is unseeded);
* ``--synthetic_table`` generates a seeded regular-solution table into
  ``--data_path`` when no table file matches the temperature tag, with
  ``--table_seed``, ``--table_diffusivity``;
* ``--precision fp64|fp32`` selects the device arithmetic of the fused step;
* ``--mode plain|fused`` selects the per-function drop-in loop or the
  device-resident fused step.

Errors: a malformed number list is an ``argparse.ArgumentTypeError`` (argparse
reports it and exits with status 2), a wrong element count a ``ValueError``.
"""
import argparse
import re
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

NUM_ELEMENTS = 5  # Fe (dependent) and four independent components

_LIST_SEP = re.compile(r"[,;]")


def number_list(count, dtype=float, shape=None):
    """An argparse type: ``count`` numbers separated by commas or semicolons (blank
    items skipped), as an array of ``dtype`` (values read as floats, then cast)."""
    def parse(text):
        items = [t.strip() for t in _LIST_SEP.split(text)]
        items = [t for t in items if t]
        if len(items) != count:
            raise argparse.ArgumentTypeError(f"{text!r} holds {len(items)} numbers; {count} are needed")
        arr = np.asarray([float(t) for t in items]).astype(dtype)
        return arr.reshape(shape) if shape is not None else arr
    return parse


# One row per flag: (group, name, type, default, extra argparse keywords).  The
# type BOOL makes the pair --name / --no_name with the given default.
BOOL = "bool"
FLAGS = [
    ("thermodynamics", "elements", str, "Fe,Mn,Ni,Co,Cu", {}),
    ("thermodynamics", "temperature", str, "873K", {}),
    ("thermodynamics", "data_path", str, "", {}),
    ("thermodynamics", "resolution", float, 0.01, {}),
    ("thermodynamics", "system_dim", int, 1, {"choices": [1, 2, 3]}),
    ("thermodynamics", "atomic_radius_tag", str, "senkov", {}),
    ("composition", "initial_composition", number_list(4), "0.1,0.2,0.2,0.5", {}),
    ("composition", "fluctuation", float, 1e-3, {}),
    ("system size", "mw", int, 100, {}),
    ("system size", "ppw", int, 16, {}),
    ("energy gradient", "kappa_value", float, 5e-16, {}),
    ("energy gradient", "kappa_i", number_list(4, shape=(1, 4)), "1,1,1,1", {}),
    ("elasticity", "include_cubic_anisotropy", BOOL, True, {}),
    ("elasticity", "direction", number_list(3, dtype=int), "1,0,0", {}),
    ("elasticity", "c11", float, None, {"help": "override c11 [GPa]"}),
    ("elasticity", "c12", float, None, {"help": "override c12 [GPa]"}),
    ("elasticity", "c44", float, None, {"help": "override c44 [GPa]"}),
    ("elasticity", "molar_volume", float, None, {"help": "override V_m [m^3/mol]"}),
    ("time scale", "total_timesteps", int, 100_000, {}),
    ("time scale", "steps_per_ctld", int, 1_000, {}),
    ("time scale", "time_linear_end_multiplier", float, 100, {}),
    ("adaptive step limits", "lower_limit", float, 1e-5, {}),
    ("adaptive step limits", "upper_limit", float, 1e-3, {}),
    ("adaptive step limits", "safety", float, 0.1, {}),
    ("save logic", "early_stage_frames", int, 100, {}),
    ("save logic", "late_burst_duration_ctld", int, 100, {}),
    ("save logic", "late_burst_frames", int, 700, {}),
    ("save logic", "late_tail_frames", int, 300, {}),
    ("save logic", "save_outputs", BOOL, True, {}),
    ("save logic", "results_root", str, "results", {}),
    ("modifier", "same_initial_configurations", BOOL, False, {}),
    ("modifier", "load_path", None, "", {}),
    ("log name", "log_name", str, None, {}),
    ("performance", "threads", int, 1, {}),
    ("performance", "fft_workers", int, 1, {}),
    ("device step", "precision", str, "fp64", {"choices": ["fp64", "fp32"]}),
    ("device step", "mode", str, "plain", {"choices": ["plain", "fused"]}),
    ("device step", "seed", int, None, {}),
    ("device step", "synthetic_table", BOOL, False, {}),
    ("device step", "table_seed", int, 0, {}),
    ("device step", "table_diffusivity", float, 1e-20, {}),
]


@dataclass
class SimulationConfig:
    elements: List[str]
    temperature: str
    data_path: str
    resolution: float
    system_dim: int
    atomic_radius_tag: str
    initial_composition: np.ndarray
    fluctuation: float
    mw: int
    ppw: int
    kappa_value: float
    kappa_i: np.ndarray
    include_cubic_anisotropy: bool
    direction: np.ndarray
    total_timesteps: int
    steps_per_ctld: int
    time_linear_end_multiplier: float
    lower_limit: float
    upper_limit: float
    safety: float
    early_stage_frames: int
    late_burst_duration_ctld: int
    late_burst_frames: int
    late_tail_frames: int
    same_initial_configurations: bool
    load_path: str
    log_name: str
    threads: int
    fft_workers: int
    # additions (defaults reproduce the reference behaviour)
    c11: Optional[float] = None
    c12: Optional[float] = None
    c44: Optional[float] = None
    molar_volume: Optional[float] = None
    seed: Optional[int] = None
    synthetic_table: bool = False
    table_seed: int = 0
    table_diffusivity: float = 1e-20
    precision: str = "fp64"
    mode: str = "plain"
    results_root: str = "results"
    save_outputs: bool = True
    extra: dict = field(default_factory=dict)


def build_parser(ncomp=4):
    """The argparse parser of FLAGS (``ncomp``: independent components, 4 here)."""
    if ncomp != NUM_ELEMENTS - 1:
        raise ValueError(f"this solver has {NUM_ELEMENTS - 1} independent components, not {ncomp}")
    parser = argparse.ArgumentParser(
        prog="simulate_spinodal",
        description="Run multi-component Cahn-Hilliard / elastic simulation (MI355X step).",
        formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    groups = {}
    for title, name, kind, default, extra in FLAGS:
        if title not in groups:
            groups[title] = parser.add_argument_group(title)
        g = groups[title]
        if kind is BOOL:
            pair = g.add_mutually_exclusive_group()
            words = name.replace("_", " ")
            pair.add_argument(f"--{name}", dest=name, action="store_true", help=f"turn {words} on")
            pair.add_argument(f"--no_{name}", dest=name, action="store_false", help=f"turn {words} off")
            parser.set_defaults(**{name: default})
        else:
            kw = dict(extra, default=default)
            if kind is not None:
                kw["type"] = kind
            g.add_argument(f"--{name}", **kw)
    return parser


def parse_args_to_config(argv=None):
    """The SimulationConfig of a command line (sys.argv when argv is None)."""
    ns = build_parser().parse_args(argv)
    elements = [e for e in (s.strip() for s in ns.elements.split(",")) if e]
    if len(elements) != NUM_ELEMENTS:
        raise ValueError(f"--elements names {len(elements)} elements ({elements}); the solver is built for "
                         f"exactly {NUM_ELEMENTS} (the first one dependent)")
    values = vars(ns)
    values["elements"] = elements
    casts = {"temperature": str, "data_path": str, "resolution": float, "system_dim": int,
             "atomic_radius_tag": str, "fluctuation": float, "mw": int, "ppw": int, "kappa_value": float,
             "include_cubic_anisotropy": bool, "total_timesteps": int, "steps_per_ctld": int,
             "time_linear_end_multiplier": float, "lower_limit": float, "upper_limit": float, "safety": float,
             "early_stage_frames": int, "late_burst_duration_ctld": int, "late_burst_frames": int,
             "late_tail_frames": int, "same_initial_configurations": bool, "load_path": str, "threads": int,
             "fft_workers": int, "synthetic_table": bool, "table_seed": int, "table_diffusivity": float,
             "save_outputs": bool}
    for key, cast in casts.items():
        values[key] = cast(values[key])
    values["initial_composition"] = np.asarray(values["initial_composition"], dtype=float)
    values["kappa_i"] = np.asarray(values["kappa_i"], dtype=float)
    values["direction"] = np.asarray(values["direction"], dtype=int)
    values["log_name"] = values["log_name"] or ""
    return SimulationConfig(**values)
