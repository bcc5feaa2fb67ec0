"""Shared test helpers: problems of the workload physics at small sizes, oracle wiring."""
import numpy as np

from oracle import pf5d_oracle as O
from phasefield5d_amd.problem import build_problem
from phasefield5d_amd.workloads import small_config


def problem(dim, n, seed=0, elastic=True, X0="0.1,0.2,0.2,0.3", kappa=5e-16, fluct=None):
    cfg = small_config(dim, n, X0=X0, kappa=kappa, seed=seed, elastic=elastic)
    if fluct is not None:
        cfg.fluctuation = fluct
    return build_problem(cfg)


def oracle_sim(pb):
    cfg = pb.cfg
    return O.OracleSim(pb.data_grid, cfg.resolution, pb.tree, pb.calphad_values, pb.dx, pb.kappa, pb.X0,
                       pb.l_vec, pb.Bhat, pb.ctld, pb.dt_cfl, cfg.lower_limit, cfg.upper_limit)


def rel_inf(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))


def developed_state(pb, amp=0.05, seed=1):
    """A decomposed-looking field: X0 + smooth random modes of amplitude ``amp`` (stays in the simplex)."""
    rng = np.random.default_rng(seed)
    shape = pb.shape
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    X = np.broadcast_to(pb.X0, tuple(shape) + (4,)).copy()
    for s in range(4):
        f = np.zeros(shape)
        for _ in range(3):
            k = [rng.integers(1, 4) for _ in shape]
            ph = rng.uniform(0, 2 * np.pi)
            f += np.cos(sum(2 * np.pi * k[d] * grids[d] / shape[d] for d in range(len(shape))) + ph)
        X[..., s] += amp / 3 * f
    return X
