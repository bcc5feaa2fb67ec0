"""Named workloads: the 1D/2D/3D single runs, the 64-run batch and the serving mix.

Common physics: Fe(dependent),Mn,Ni,Co,Cu; T = 873 K; synthetic regular-solution
table at step 0.025 (D = 1e-20 m^2/s); elasticity on with c11/c12/c44 =
240/150/120 GPa, senkov Vegard strains, V_m = 7.1e-6 m^3/mol; fluctuation 1e-3
seeded; steps_per_ctld = 1000.
"""
import numpy as np

from .solver.config import parse_args_to_config

COMMON = ["--elements", "Fe,Mn,Ni,Co,Cu", "--temperature", "873K", "--resolution", "0.025",
          "--c11", "240", "--c12", "150", "--c44", "120", "--molar_volume", "7.1e-6",
          "--steps_per_ctld", "1000", "--fluctuation", "1e-3", "--kappa_i", "1,1,1,1",
          "--include_cubic_anisotropy", "--synthetic_table"]

WORKLOADS = {
    # 1D periodic, 160 points, direction 1,0,0, 1e4 steps (CPU-runnable case)
    "1d_160": ["--system_dim", "1", "--mw", "20", "--ppw", "8", "--direction", "1,0,0",
               "--initial_composition", "0.1,0.2,0.2,0.3", "--kappa_value", "5e-16",
               "--total_timesteps", "10000"],
    # 2D periodic 256^2, plane normal 0,0,1, 2e4 steps (plain FP64 drop-in)
    "2d_256": ["--system_dim", "2", "--mw", "16", "--ppw", "16", "--direction", "0,0,1",
               "--initial_composition", "0.1,0.2,0.2,0.3", "--kappa_value", "7.6e-15",
               "--total_timesteps", "20000"],
    # 3D periodic 128^3, 5e3 steps (precision form)
    "3d_128": ["--system_dim", "3", "--mw", "8", "--ppw", "16", "--direction", "1,0,0",
               "--initial_composition", "0.1,0.2,0.3,0.2", "--kappa_value", "5e-16",
               "--total_timesteps", "5000"],
}


def workload_config(name, seed=0, table_seed=0, extra=()):
    cfg = parse_args_to_config(COMMON + WORKLOADS[name] + ["--seed", str(seed), "--table_seed", str(table_seed)]
                               + list(extra))
    return cfg


def small_config(dim, n, X0="0.1,0.2,0.2,0.3", kappa=5e-16, seed=0, elastic=True, ppw=None):
    """A reduced-size run of the same physics (parity tests): n cells per axis."""
    ppw = ppw or (8 if n % 8 == 0 else n)
    mw = n // ppw
    args = COMMON + ["--system_dim", str(dim), "--mw", str(mw), "--ppw", str(ppw), "--direction",
                     "1,0,0" if dim != 2 else "0,0,1", "--initial_composition", X0, "--kappa_value", str(kappa),
                     "--seed", str(seed)]
    if not elastic:
        args = [a for a in args if a != "--include_cubic_anisotropy"] + ["--no_include_cubic_anisotropy"]
    return parse_args_to_config(args)
