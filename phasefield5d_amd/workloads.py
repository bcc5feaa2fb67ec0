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


TABLE_X0 = "0.1,0.2,0.2,0.3"   # composition the per-temperature tables are drawn spinodal at
TABLE_SEEDS = {873: 873, 1000: 1000}


def _mix_config(dim, side, X0, kappa, T, seed, ppw=16):
    mw = max(1, side // ppw)
    args = COMMON + ["--system_dim", str(dim), "--mw", str(mw), "--ppw", str(side // mw),
                     "--direction", "0,0,1" if dim == 2 else "1,0,0",
                     "--initial_composition", ",".join(f"{v:.6f}" for v in X0), "--kappa_value", f"{kappa:.6e}",
                     "--temperature", f"{T}K", "--seed", str(seed), "--table_seed", str(TABLE_SEEDS[T])]
    return parse_args_to_config(args)


def draw_spinodal_composition(rng, T, max_sum=0.85, tries=200):
    """X0 ~ Dirichlet(2,2,2,2,2) restricted to sum(X) <= max_sum and to the spinodal
    region of the T table (compute_initial_kinetics must find a growing mode)."""
    from .problem import build_problem
    for _ in range(tries):
        x5 = rng.dirichlet([2, 2, 2, 2, 2])
        X0 = x5[1:]
        if X0.sum() > max_sum:
            continue
        cfg = _mix_config(1, 32, X0, 1e-15, T, 0)
        tab = table_for(T)
        try:
            build_problem(cfg, table=tab)
        except ValueError:
            continue
        return X0
    raise RuntimeError("no spinodal composition drawn")


_TABLES = {}


def table_for(T):
    """The seeded synthetic table of temperature T (spinodal at TABLE_X0), rows as the loader returns them."""
    if T not in _TABLES:
        from .thermodynamics.synthetic import make_synthetic_table
        X0 = np.array([float(v) for v in TABLE_X0.split(",")])
        _TABLES[T] = make_synthetic_table(X0, T=float(T), seed=TABLE_SEEDS[T], step=0.025)
    return _TABLES[T]


def batch_mix(n2d=32, n3d=32, seed=0, sides2d=(128, 192, 256, 384, 512), sides3d=(64, 96, 128)):
    """The 64-run batch: 2D sides from sides2d, 3D sides from sides3d (uniform draws),
    X0 from a spinodal-restricted Dirichlet(2,...,2), kappa log-uniform in [1e-16, 1e-14],
    T in {873 K, 1000 K} with one seeded table each.  Returns (config, table) pairs."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n2d + n3d):
        dim = 2 if i < n2d else 3
        side = int(rng.choice(sides2d if dim == 2 else sides3d))
        T = int(rng.choice([873, 1000]))
        X0 = draw_spinodal_composition(rng, T)
        kappa = float(10 ** rng.uniform(-16, -14))
        out.append((_mix_config(dim, side, X0, kappa, T, seed * 1000 + i), table_for(T)))
    return out


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


def developed_field(pb, amp=0.05, seed=1):
    """A composition field that looks decomposed: X0 plus, per component, three
    smooth periodic modes (wavenumbers 1-3 per axis, random phases) of total
    amplitude ``amp``.  Used to time the step on fields like those later in a run,
    where neighbouring cells no longer share a table hypercube."""
    rng = np.random.default_rng(seed)
    shape = tuple(pb.shape)
    axes = [np.arange(n) for n in shape]
    X = np.broadcast_to(pb.X0, shape + (4,)).copy()
    for s in range(4):
        f = np.zeros(shape)
        for _ in range(3):
            k = [int(rng.integers(1, 4)) for _ in shape]
            ph = rng.uniform(0, 2 * np.pi)
            arg = ph
            for d in range(len(shape)):
                sh = [1] * len(shape)
                sh[d] = shape[d]
                arg = arg + (2 * np.pi * k[d] / shape[d]) * axes[d].reshape(sh)
            f += np.cos(arg)
        X[..., s] += amp / 3 * f
    return X

