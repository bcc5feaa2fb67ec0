"""Seeded synthetic regular-solution CALPHAD tables.

No ThermoCalc table ships with the reference, so every workload runs on a
synthetic table with the layout of data/FeMnNiCoCu_fcc/DATA_FORMAT.md:19-49:

* nodes: X_i in {0, h, 2h, ...} for the four independent components with
  sum(X) <= 1 (41^4 box, 135,751 simplex nodes at h = 0.025);
* Gm.x(i) = dG/dX_i with X_Fe = 1 - sum(X) dependent, where
  G = RT sum_e X_e ln X_e + sum_{e<f} Omega_ef X_e X_f over all five elements
  (ten pairs).  With g_e = RT (ln X_e + 1) + sum_{f != e} Omega_ef X_f this is
  g_i - g_Fe.  X_e is floored at ``eps_log`` inside the logarithm only;
* Mob(E) = D / RT for all five elements, so the dyadic face mobility of the
  step reduces to M_ij = (X_i delta_ij - X_i X_j) D / RT;
* Omega_ef ~ U(omega_lo, omega_hi) J/mol from a seeded generator, redrawn until
  the finite-difference Hessian at X0 (the one the kinetics setup builds,
  optionally plus an elastic matrix) has a negative eigenvalue.
"""
from dataclasses import dataclass

import numpy as np
import pandas as pd

R_GAS = 8.314462618
ELEMENTS = ["Fe", "Mn", "Ni", "Co", "Cu"]


def simplex_nodes(step):
    """All grid nodes (X_Mn, X_Ni, X_Co, X_Cu) with sum <= 1, index order i0 slowest."""
    num = int(round(1.0 / step)) + 1
    idx = np.indices((num,) * 4).reshape(4, -1).T
    idx = idx[idx.sum(axis=1) <= num - 1]
    return idx, idx * step


def _omega_matrix(pairs):
    om = np.zeros((5, 5))
    k = 0
    for e in range(5):
        for f in range(e + 1, 5):
            om[e, f] = om[f, e] = pairs[k]
            k += 1
    return om


def regular_solution_mu(X4, T, omega, eps_log=1e-8):
    """dG/dX_i (i = Mn..Cu) for rows of independent mole fractions ``X4`` (n, 4)."""
    X4 = np.atleast_2d(np.asarray(X4, dtype=float))
    X5 = np.concatenate([1.0 - X4.sum(axis=1, keepdims=True), X4], axis=1)
    RT = R_GAS * T
    g = RT * (np.log(np.maximum(X5, eps_log)) + 1.0) + X5 @ omega   # omega has zero diagonal
    return g[:, 1:] - g[:, :1]


@dataclass
class SyntheticTable:
    temperature: float
    step: float
    omega: np.ndarray          # (5, 5) J/mol, zero diagonal
    diffusivity: float
    rows: np.ndarray           # (n_nodes, 13): 4 X, 4 Gm.x, 5 Mob
    seed: int
    draws: int

    @property
    def mobility(self):
        return self.diffusivity / (R_GAS * self.temperature)

    def dataframe(self):
        cols = ([f"x({e})" for e in ELEMENTS[1:]] + [f"Gm.x({e})" for e in ELEMENTS[1:]]
                + [f"Mob({e})" for e in ELEMENTS])
        return pd.DataFrame(self.rows, columns=cols)

    def renamed_dataframe(self):
        """Columns renamed as the TSV loader returns them (X_E, Gx_E, Mob_E)."""
        df = self.dataframe()
        df.columns = ([f"X_{e}" for e in ELEMENTS[1:]] + [f"Gx_{e}" for e in ELEMENTS[1:]]
                      + [f"Mob_{e}" for e in ELEMENTS])
        return df


def _fd_hessian(X0, step, T, omega):
    """Central-difference Hessian on grid nodes, as the kinetics setup evaluates it."""
    X0 = np.asarray(X0, dtype=float)
    pts_p = X0[None, :] + np.eye(4) * step
    pts_m = X0[None, :] - np.eye(4) * step
    mu_p = regular_solution_mu(pts_p, T, omega)
    mu_m = regular_solution_mu(pts_m, T, omega)
    return (mu_p - mu_m).T / (2.0 * step)


def make_synthetic_table(X0, T=873.0, seed=0, step=0.025, diffusivity=1e-20,
                         omega_range=(-10e3, 25e3), elastic_matrix=None, max_draws=1000):
    """Draw Omega until the Hessian (+ elastic matrix) at X0 is indefinite; build the rows."""
    rng = np.random.default_rng(seed)
    X0 = np.asarray(X0, dtype=float)
    for draw in range(1, max_draws + 1):
        omega = _omega_matrix(rng.uniform(omega_range[0], omega_range[1], size=10))
        H = _fd_hessian(X0, step, T, omega)
        if elastic_matrix is not None:
            H = H + elastic_matrix
        if np.linalg.eigvalsh(0.5 * (H + H.T))[0] < 0:
            break
    else:
        raise RuntimeError("could not draw a spinodal regular-solution table")
    _, X = simplex_nodes(step)
    mu = regular_solution_mu(X, T, omega)
    mob = np.full((len(X), 5), diffusivity / (R_GAS * T))
    rows = np.concatenate([X, mu, mob], axis=1)
    return SyntheticTable(float(T), float(step), omega, float(diffusivity), rows, int(seed), draw)


def write_synthetic_table(table, directory, name=None):
    """Write ``table`` as ``FeMnNiCoCu_fcc_<T>K.txt`` (tab separated) and return the path."""
    import os
    os.makedirs(directory, exist_ok=True)
    name = name or f"FeMnNiCoCu_fcc_{int(round(table.temperature))}K.txt"
    path = os.path.join(directory, name)
    table.dataframe().to_csv(path, sep="\t", index=False, float_format="%.17g", lineterminator="\n")
    return path
