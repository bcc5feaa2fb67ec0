"""Once-per-run setup of a simulation (examples/simulate.py:59-221).

``build_problem(cfg)`` performs the reference's setup sequence on the host:
material constants, CALPHAD table -> dense grid + KD-tree, reference-point
potentials (``hessian_max`` is max|mu(X0)|, as simulate.py:86 computes it),
the rank-1 elastic coupling, CTLD and the dominant wavenumber, dx / N / dt0,
the CFL bound, the Fourier grid and B(k), and the seeded initial field.
"""
import os
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .elasticity.constants import load_elastic_constants, load_material_constants
from .elasticity.energies import calculate_molar_volume
from .solver.elastic import (build_khachaturyan_kernel, calculate_linear_elastic_coupling_matrix,
                             get_elastic_matrix, rank_one_vector)
from .solver.interpolation import build_4d_grid, build_calphad_kdtree, interpolate_points_host
from .solver.system import (apply_fluctuations, build_fourier_grid_1d_along_direction,
                            build_fourier_grid_2d_in_plane, build_fourier_grid_3d, cahn_hilliard_cfl_dt,
                            compute_initial_kinetics, get_system_dimensions, get_time_increment)
from .thermodynamics.io import find_table_file, load_calphad_dataframe
from .thermodynamics.synthetic import make_synthetic_table, write_synthetic_table


def temperature_kelvin(tag):
    s = str(tag).strip().lower().rstrip("k")
    return float(s)


@dataclass
class Problem:
    cfg: object
    rows: np.ndarray
    data_grid: np.ndarray
    tree: object
    calphad_values: np.ndarray
    hessian_max: float
    mobility_max: float
    cij: tuple
    molar_volume: float
    Lambda: Optional[np.ndarray]
    l_vec: Optional[np.ndarray]
    elastic_matrix: Optional[np.ndarray]
    ctld: float
    k_max: float
    dt0: float
    dx: float
    shape: tuple
    dt_cfl: float
    Bhat: Optional[np.ndarray]
    kappa: np.ndarray
    X_init: np.ndarray
    extra: dict = field(default_factory=dict)

    @property
    def ncell(self):
        return int(np.prod(self.shape))

    @property
    def X0(self):
        return np.asarray(self.cfg.initial_composition, dtype=np.float64)


_SYN_CACHE = {}


def synthetic_rows(X0, T, seed, step, diffusivity):
    """In-memory synthetic table rows (cached per parameter set)."""
    key = (X0, T, seed, step, diffusivity)
    if key not in _SYN_CACHE:
        tab = make_synthetic_table(np.asarray(X0), T=T, seed=seed, step=step, diffusivity=diffusivity)
        _SYN_CACHE[key] = tab.renamed_dataframe().to_numpy()
    return _SYN_CACHE[key]


def table_rows(cfg, table=None, verbose=False):
    """Rows (n, 13) of the CALPHAD table for ``cfg``: given, on disk, or synthetic."""
    if table is not None:
        # the rows array of a SyntheticTable is stable, so its dense grid is cached once
        return table.rows if hasattr(table, "rows") else np.asarray(table)
    if not cfg.data_path and cfg.synthetic_table:
        return synthetic_rows(tuple(np.asarray(cfg.initial_composition, dtype=float)),
                              temperature_kelvin(cfg.temperature), cfg.table_seed, cfg.resolution,
                              cfg.table_diffusivity)
    path = cfg.data_path or os.path.join(os.getcwd(), "data", "FeMnNiCoCu_fcc")
    try:
        find_table_file(cfg.temperature, path)
    except (FileNotFoundError, OSError):
        if not cfg.synthetic_table:
            raise
        T = temperature_kelvin(cfg.temperature)
        tab = make_synthetic_table(cfg.initial_composition, T=T, seed=cfg.table_seed, step=cfg.resolution,
                                   diffusivity=cfg.table_diffusivity)
        write_synthetic_table(tab, path)
        if verbose:
            print(f"wrote synthetic table T={T}K seed={cfg.table_seed} ({tab.draws} draws) to {path}")
    return load_calphad_dataframe(cfg.temperature, path=path, elements=cfg.elements).to_numpy()


_GRID_CACHE = {}


def grid_and_tree(rows, resolution):
    """Dense grid + KD-tree of a table, cached while the rows array is alive."""
    key = (id(rows), float(resolution))
    hit = _GRID_CACHE.get(key)
    if hit is not None and hit[0] is rows:
        return hit[1], hit[2], hit[3]
    _, grid = build_4d_grid(rows, resolution)
    tree, values = build_calphad_kdtree(rows)
    if len(_GRID_CACHE) > 8:
        _GRID_CACHE.clear()
    _GRID_CACHE[key] = (rows, grid, tree, values)
    return grid, tree, values


def build_problem(cfg, table=None, verbose=False, X_init=None):
    alloy, ri, mu, nu, vi, qi = load_material_constants(cfg.atomic_radius_tag, alloy=cfg.elements)
    override = (cfg.c11, cfg.c12, cfg.c44) if cfg.c11 is not None else None
    c11, c12, c44 = load_elastic_constants(override=override)
    rows = table_rows(cfg, table, verbose)
    grid, tree, values = grid_and_tree(rows, cfg.resolution)
    X0 = np.asarray(cfg.initial_composition, dtype=float)
    ref = interpolate_points_host(X0[None, :], grid, cfg.resolution, tree, values)[0]
    hessian_max = float(np.abs(ref[:4]).max())
    mobility_max = float(ref[4:].max())
    vm = cfg.molar_volume if cfg.molar_volume is not None else calculate_molar_volume(X0, vi)
    Lam = l_vec = emat = None
    if cfg.include_cubic_anisotropy:
        Lam = calculate_linear_elastic_coupling_matrix(X0, ri) * vm
        emat = get_elastic_matrix(X0, vi, ri, c11, c12, c44, cfg.direction, molar_volume=vm)
    ctld, kmax = compute_initial_kinetics(X0, cfg.resolution, cfg.kappa_value, cfg.kappa_i, grid, tree, values,
                                          elastic_matrix=emat, n_k=10_000, verbose=verbose)
    dt0 = get_time_increment(ctld, cfg.steps_per_ctld)
    _, dx, _, shape = get_system_dimensions(kmax, ppw=cfg.ppw, mw=cfg.mw, dim=cfg.system_dim)
    dt_cfl = cahn_hilliard_cfl_dt(dx, mobility_max, hessian_max, cfg.kappa_value, cfg.safety)
    Bhat = None
    if cfg.include_cubic_anisotropy:
        if cfg.system_dim == 1:
            kg, kn = build_fourier_grid_1d_along_direction(shape[0], dx, cfg.direction)
        elif cfg.system_dim == 2:
            kg, kn = build_fourier_grid_2d_in_plane(shape[0], shape[1], dx, cfg.direction)
        else:
            kg, kn = build_fourier_grid_3d(shape[0], shape[1], shape[2], dx)
        Bhat = build_khachaturyan_kernel(kg, kn, c11, c12, c44)
        l_vec, rank1 = rank_one_vector(Lam)
        if not rank1:
            raise NotImplementedError("non-rank-1 elastic coupling")
    if X_init is None:
        X_init = np.asarray(apply_fluctuations(X0, shape, cfg.fluctuation, seed=cfg.seed), dtype=np.float64)
    kappa = np.asarray(cfg.kappa_i, dtype=float).reshape(1, 4) * cfg.kappa_value
    return Problem(cfg=cfg, rows=rows, data_grid=grid, tree=tree, calphad_values=values,
                   hessian_max=hessian_max, mobility_max=mobility_max, cij=(c11, c12, c44), molar_volume=vm,
                   Lambda=Lam, l_vec=l_vec, elastic_matrix=emat, ctld=ctld, k_max=kmax, dt0=dt0, dx=dx,
                   shape=tuple(shape), dt_cfl=dt_cfl, Bhat=Bhat, kappa=kappa, X_init=X_init)
