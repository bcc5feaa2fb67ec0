"""Run configuration: dataclass + argparse parser.

Keeps every flag of phasefield5d/solver/config.py:74-143 with the same names,
defaults and parsing (``--x/--no_x`` boolean pairs, comma lists, ``direction``
cast to int, exactly five elements).  Added flags (all defaulting to the
reference behaviour):

* ``--c11/--c12/--c44`` (GPa) override the hard-coded cubic constants;
* ``--molar_volume`` (m^3/mol) overrides the linear-mixing molar volume;
* ``--seed`` seeds the initial fluctuation draw (the reference draw is unseeded);
* ``--synthetic_table`` generates a seeded regular-solution table into
  ``--data_path`` when no table file matches the temperature tag, with
  ``--table_seed``, ``--table_diffusivity``;
* ``--precision fp64|fp32`` selects the device arithmetic of the fused step;
* ``--mode plain|fused`` selects the per-function drop-in loop or the
  device-resident fused step.
"""
import argparse
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np


def parse_float_list(s, *, expected_len=None):
    vals = [float(x) for x in s.replace(";", ",").split(",") if x.strip() != ""]
    if expected_len is not None and len(vals) != expected_len:
        raise argparse.ArgumentTypeError(f"Expected {expected_len} values, got {len(vals)}")
    return vals


def composition_type(s, *, ncomp):
    return np.array(parse_float_list(s, expected_len=ncomp), dtype=float)


def add_bool_flag(parser, name, default=False, help_on="", help_off=""):
    dest = name.replace("-", "_")
    grp = parser.add_mutually_exclusive_group()
    grp.add_argument(f"--{name}", dest=dest, action="store_true",
                     help=help_on or f"Enable {name.replace('-', ' ')}.")
    grp.add_argument(f"--no_{name}", dest=dest, action="store_false",
                     help=help_off or f"Disable {name.replace('-', ' ')}.")
    parser.set_defaults(**{dest: default})


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
    p = argparse.ArgumentParser(
        prog="simulate_spinodal",
        description="Run multi-component Cahn-Hilliard / elastic simulation (MI355X step).",
        formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    th = p.add_argument_group("thermodynamics")
    th.add_argument("--elements", type=str, default="Fe,Mn,Ni,Co,Cu")
    th.add_argument("--temperature", type=str, default="873K")
    th.add_argument("--data_path", type=str, default="")
    th.add_argument("--resolution", type=float, default=0.01)
    th.add_argument("--system_dim", type=int, choices=[1, 2, 3], default=1)
    th.add_argument("--atomic_radius_tag", type=str, default="senkov")
    co = p.add_argument_group("composition")
    co.add_argument("--initial_composition", type=lambda s: composition_type(s, ncomp=4),
                    default="0.1,0.2,0.2,0.5")
    co.add_argument("--fluctuation", type=float, default=1e-3)
    sz = p.add_argument_group("system size")
    sz.add_argument("--mw", type=int, default=100)
    sz.add_argument("--ppw", type=int, default=16)
    ka = p.add_argument_group("energy gradient")
    ka.add_argument("--kappa_value", type=float, default=5e-16)
    ka.add_argument("--kappa_i", type=lambda s: np.array(parse_float_list(s, expected_len=4),
                                                         dtype=float).reshape(1, 4), default="1,1,1,1")
    el = p.add_argument_group("elasticity")
    add_bool_flag(el, "include_cubic_anisotropy", default=True)
    el.add_argument("--direction", type=lambda s: np.array(parse_float_list(s, expected_len=3), dtype=int),
                    default="1,0,0")
    el.add_argument("--c11", type=float, default=None, help="override c11 [GPa]")
    el.add_argument("--c12", type=float, default=None, help="override c12 [GPa]")
    el.add_argument("--c44", type=float, default=None, help="override c44 [GPa]")
    el.add_argument("--molar_volume", type=float, default=None, help="override V_m [m^3/mol]")
    ts = p.add_argument_group("time scale")
    ts.add_argument("--total_timesteps", type=int, default=100_000)
    ts.add_argument("--steps_per_ctld", type=int, default=1_000)
    ts.add_argument("--time_linear_end_multiplier", type=float, default=100)
    ad = p.add_argument_group("adaptive step limits")
    ad.add_argument("--lower_limit", type=float, default=1e-5)
    ad.add_argument("--upper_limit", type=float, default=1e-3)
    ad.add_argument("--safety", type=float, default=0.1)
    sv = p.add_argument_group("save logic")
    sv.add_argument("--early_stage_frames", type=int, default=100)
    sv.add_argument("--late_burst_duration_ctld", type=int, default=100)
    sv.add_argument("--late_burst_frames", type=int, default=700)
    sv.add_argument("--late_tail_frames", type=int, default=300)
    add_bool_flag(sv, "save_outputs", default=True)
    sv.add_argument("--results_root", type=str, default="results")
    md = p.add_argument_group("modifier")
    add_bool_flag(md, "same_initial_configurations", default=False)
    md.add_argument("--load_path", default="")
    lg = p.add_argument_group("log name")
    lg.add_argument("--log_name", type=str, default=None)
    pf = p.add_argument_group("performance")
    pf.add_argument("--threads", type=int, default=1)
    pf.add_argument("--fft_workers", type=int, default=1)
    dv = p.add_argument_group("device step")
    dv.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    dv.add_argument("--mode", choices=["plain", "fused"], default="plain")
    dv.add_argument("--seed", type=int, default=None)
    add_bool_flag(dv, "synthetic_table", default=False)
    dv.add_argument("--table_seed", type=int, default=0)
    dv.add_argument("--table_diffusivity", type=float, default=1e-20)
    return p


def parse_args_to_config(argv=None):
    a = build_parser(ncomp=4).parse_args(argv)
    elements = [e.strip() for e in a.elements.split(",") if e.strip()]
    if len(elements) != 5:
        raise ValueError(f"--elements must list exactly 5 elements (got {len(elements)}: {elements}). "
                         "phasefield5d is a 5-component solver; the number of components is fixed.")
    return SimulationConfig(
        elements=elements, temperature=str(a.temperature), data_path=str(a.data_path),
        resolution=float(a.resolution), system_dim=int(a.system_dim),
        atomic_radius_tag=str(a.atomic_radius_tag),
        initial_composition=np.asarray(a.initial_composition, dtype=float),
        fluctuation=float(a.fluctuation), mw=int(a.mw), ppw=int(a.ppw),
        kappa_value=float(a.kappa_value), kappa_i=np.asarray(a.kappa_i, dtype=float),
        include_cubic_anisotropy=bool(a.include_cubic_anisotropy),
        direction=np.asarray(a.direction, dtype=int), total_timesteps=int(a.total_timesteps),
        steps_per_ctld=int(a.steps_per_ctld),
        time_linear_end_multiplier=float(a.time_linear_end_multiplier),
        lower_limit=float(a.lower_limit), upper_limit=float(a.upper_limit), safety=float(a.safety),
        early_stage_frames=int(a.early_stage_frames),
        late_burst_duration_ctld=int(a.late_burst_duration_ctld),
        late_burst_frames=int(a.late_burst_frames), late_tail_frames=int(a.late_tail_frames),
        same_initial_configurations=bool(a.same_initial_configurations),
        load_path=str(a.load_path), log_name=a.log_name or "", threads=int(a.threads),
        fft_workers=int(a.fft_workers), c11=a.c11, c12=a.c12, c44=a.c44,
        molar_volume=a.molar_volume, seed=a.seed, synthetic_table=bool(a.synthetic_table),
        table_seed=int(a.table_seed), table_diffusivity=float(a.table_diffusivity),
        precision=a.precision, mode=a.mode, results_root=a.results_root,
        save_outputs=bool(a.save_outputs))
