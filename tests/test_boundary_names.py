"""The import swap of the reference driver (examples/simulate.py:15-58) resolves in
phasefield5d_amd, and the host-side helpers it calls accept the reference's
keyword arguments (make_path at simulate.py:160-171, should_save at :291-297).
CPU only: nothing here touches the GPU."""
import importlib
import json
import os

import numpy as np
import pytest

from oracle import pf5d_oracle as O

# module -> names, as examples/simulate.py:15-58 takes them from phasefield5d
DRIVER_IMPORTS = {
    "solver.config": ["parse_args_to_config"],
    "thermodynamics.io": ["load_calphad_dataframe"],
    "elasticity.constants": ["load_material_constants", "load_elastic_constants"],
    "elasticity.energies": ["calculate_molar_volume"],
    "solver.interpolation": ["build_4d_grid", "build_calphad_kdtree", "interpolator_nb"],
    "solver.system": ["apply_fluctuations", "out_of_bounds", "apply_adaptive_time", "get_time_increment",
                      "get_system_dimensions", "cahn_hilliard_cfl_dt", "build_fourier_grid_1d_along_direction",
                      "build_fourier_grid_2d_in_plane", "build_fourier_grid_3d", "compute_initial_kinetics"],
    "solver.operators": ["calculate_laplacian", "calculate_gradients_pm", "abs_max", "clamp_simplex"],
    "solver.fluxes": ["get_flux_function", "get_divergence_from_mu_function"],
    "solver.elastic": ["build_khachaturyan_kernel", "calculate_linear_elastic_coupling_matrix",
                       "make_elastic_updater", "get_elastic_matrix"],
    "solver.io": ["should_save", "make_path", "save_snapshot_figure", "save_current_state", "init_traces",
                  "update_traces", "finalize_traces"],
    "utils.diagnostics": ["print_time_diagnositcs", "debugging_update"],
}

# phasefield5d/solver/__init__.py:16-70 re-exports (eager and lazy)
SOLVER_REEXPORTS = [
    "SimulationConfig", "parse_args_to_config", "out_of_bounds", "apply_fluctuations", "apply_adaptive_time",
    "get_time_increment", "get_system_dimensions", "cahn_hilliard_cfl_dt", "snap_to_grid_floor",
    "build_fourier_grid", "build_fourier_grid_1d_along_direction", "build_fourier_grid_2d_in_plane",
    "build_fourier_grid_3d", "solve_elastic_kinetic_eigenproblem", "compute_initial_kinetics",
    "make_elastic_updater", "build_khachaturyan_kernel", "calculate_elastic_potential",
    "calculate_linear_elastic_coupling_matrix", "get_elastic_matrix", "should_save", "make_path",
    "save_current_state", "save_snapshot_figure", "init_traces", "update_traces", "finalize_traces",
    "get_model_tag", "iter_npz_states", "parse_run_name", "npz_to_vtk", "npz_to_vti", "batch_npz_to_vti",
    "calculate_laplacian", "calculate_gradients_pm", "get_flux_function", "calculate_mobility_matrix_pm",
    "build_4d_grid", "build_calphad_kdtree", "interpolator_nb",
]


@pytest.mark.parametrize("mod", sorted(DRIVER_IMPORTS))
def test_driver_imports_resolve(mod):
    m = importlib.import_module(f"phasefield5d_amd.{mod}")
    missing = [n for n in DRIVER_IMPORTS[mod] if not callable(getattr(m, n, None))]
    assert not missing, f"phasefield5d_amd.{mod} lacks {missing}"


def test_solver_package_reexports():
    import phasefield5d_amd.solver as S
    assert not [n for n in SOLVER_REEXPORTS if not hasattr(S, n)]


def _cfg(tmp_path):
    from phasefield5d_amd.workloads import small_config
    cfg = small_config(2, 16, seed=0)
    cfg.results_root = str(tmp_path)
    return cfg


def test_make_path_reference_keywords(tmp_path):
    from phasefield5d_amd.solver.io import make_path
    cfg = _cfg(tmp_path)
    path = make_path(cfg=cfg, hessian_max=1234.5, mobility_max=1.38e-24, number_of_cells=[16, 16],
                     cell_size=2e-9, system_length=3.2e-8, ctld=12.5, wavenumber_max=1.9e8,
                     wavelength_max=3.3e-8, time_increment=0.02)
    assert os.path.isdir(os.path.join(path, "data")) and os.path.isdir(os.path.join(path, "snapshots"))
    name = os.path.basename(path)
    assert name.startswith("cells16x16_dx2e-09_fluc") and "_dt2e-02s_" in name
    assert os.path.basename(os.path.dirname(path)) == "elastic_cubic_normal001"
    meta = json.load(open(os.path.join(path, "metadata.json")))
    assert meta["hessian_max"] == 1234.5 and meta["ctld"] == 12.5 and meta["number_of_cells"] == [16, 16]
    assert meta["time_increment"] == 0.02 and meta["run_dir"] == path


def test_should_save_reference_keywords_and_state():
    """Early phase: a data save at the first step whose time reaches each multiple of
    ctld / early_frames, a snapshot at each multiple of ctld / (early_frames // 10);
    step 0 and the last step always save, and so does step 1 (the next save time starts
    at 0 and step 0 does not advance it); step 0 restarts the schedule."""
    from phasefield5d_amd.solver.io import should_save
    ctld, dt, total = 2.0, 0.0013, 1000
    kw = dict(early_frames=40, burst_duration_ctld=100.0, burst_frames=800, tail_dt_ctld=10.0, snapshot_factor=10)
    for _ in range(2):  # a second run must start afresh at step 0
        data, snap = [], []
        for step in range(total):
            t_after = (step + 1) * dt
            d, s = should_save(step, ctld, t_after, total, **kw)
            if d:
                data.append(step)
            if s:
                snap.append(step)
        assert data[0] == 0 and snap[0] == 0 and data[-1] == total - 1 and snap[-1] == total - 1
        # t_after = (step+1) dt reaches k * ctld/40 at step ceil(k ctld / 40 / dt) - 1
        t_end = total * dt
        exp_d = sorted({0, 1, total - 1} | {int(np.ceil(k * (ctld / 40) / dt - 1e-12)) - 1
                                          for k in range(1, int(t_end / (ctld / 40)) + 1)})
        exp_s = sorted({0, 1, total - 1} | {int(np.ceil(k * (ctld / 4) / dt - 1e-12)) - 1
                                          for k in range(1, int(t_end / (ctld / 4)) + 1)})
        assert data == [s for s in exp_d if s < total] and snap == [s for s in exp_s if s < total]
        assert isinstance(should_save._state, dict) and should_save._state["next_data"] > t_end - ctld / 40
    # every-N-step saves (io.py:104-107)
    should_save(0, 1e9, 0.0, 10 ** 6)
    assert should_save(1, 1e9, 0.5, 10 ** 6) == (True, True)  # next save times start at 0
    assert should_save(5, 1e9, 1.0, 10 ** 6, save_every_n_steps_data=5, save_every_n_steps_snap=None) == (True, False)
    assert should_save(6, 1e9, 1.0, 10 ** 6, save_every_n_steps_data=5, save_every_n_steps_snap=None) == (False, False)


def test_mobility_matrix_pm_is_the_face_flux_operator():
    """M(face) g equals the oracle's face flux (itself pinned against closed forms)."""
    from phasefield5d_amd.solver.mobility import calculate_mobility_matrix_pm
    rng = np.random.default_rng(3)
    X = rng.uniform(0.05, 0.2, size=(6, 5, 4))
    Mob = rng.uniform(0.5, 2.0, size=(6, 5, 5)) * 1e-24
    g = rng.normal(size=(6, 5, 4))
    Mp, Mm = calculate_mobility_matrix_pm(X, Mob)
    assert Mp.shape == (2, 6, 5, 4, 4) and Mm.shape == (2, 6, 5, 4, 4)
    for d in range(2):
        fp = O.face_flux(X, np.roll(X, -1, axis=d), Mob, np.roll(Mob, -1, axis=d), g)
        fm = O.face_flux(X, np.roll(X, 1, axis=d), Mob, np.roll(Mob, 1, axis=d), g)
        np.testing.assert_allclose(np.einsum("...ab,...b->...a", Mp[d], g), fp, rtol=1e-12, atol=1e-40)
        np.testing.assert_allclose(np.einsum("...ab,...b->...a", Mm[d], g), fm, rtol=1e-12, atol=1e-40)


def test_snap_to_grid_floor():
    from phasefield5d_amd.solver.system import snap_to_grid_floor
    c = np.array([[0.1, 0.2, 0.3, 0.2],          # on nodes: unchanged
                  [0.1049, 0.2, -0.3, 0.2],       # floor, clip below 0
                  [0.6, 0.6, 0.0, 0.0],           # sum 1.2: scaled to 0.5, 0.5
                  [1.7, 0.0, 0.0, 0.0]])          # clip above 1
    out = snap_to_grid_floor(c, step=0.01)
    np.testing.assert_array_equal(out, [[0.1, 0.2, 0.3, 0.2], [0.1, 0.2, 0.0, 0.2], [0.5, 0.5, 0.0, 0.0],
                                        [1.0, 0.0, 0.0, 0.0]])
    assert snap_to_grid_floor(np.array([[0.07, 0.0, 0.0, 0.0]]), step=0.025)[0, 0] == 0.05
