"""phasefield5d_amd.solver -- reference-named step functions backed by libpf5d.

Host setup helpers import eagerly; device-backed names need the built library
and a GPU only when called.
"""
from .config import SimulationConfig, parse_args_to_config, build_parser  # noqa: F401
from .system import (  # noqa: F401
    out_of_bounds, apply_fluctuations, apply_adaptive_time, get_time_increment, get_system_dimensions,
    cahn_hilliard_cfl_dt, build_fourier_grid, build_fourier_grid_1d_along_direction,
    build_fourier_grid_2d_in_plane, build_fourier_grid_3d, solve_elastic_kinetic_eigenproblem,
    compute_initial_kinetics, snap_to_grid_floor,
)
from .elastic import (  # noqa: F401
    make_elastic_updater, build_khachaturyan_kernel, calculate_linear_elastic_coupling_matrix, get_elastic_matrix,
    calculate_elastic_potential,
)
from .interpolation import build_4d_grid, build_calphad_kdtree, interpolator_nb  # noqa: F401
from .operators import calculate_laplacian, calculate_gradients_pm, abs_max, clamp_simplex  # noqa: F401
from .fluxes import get_divergence_from_mu_function, get_flux_function  # noqa: F401
from .io import (  # noqa: F401
    SaveSchedule, should_save, make_path, make_problem_path, save_current_state, save_snapshot_figure, init_traces,
    update_traces, finalize_traces, get_model_tag,
)
from .mobility import calculate_mobility_matrix_pm  # noqa: F401
from .post_process import iter_npz_states, parse_run_name, npz_to_vtk, npz_to_vti, batch_npz_to_vti  # noqa: F401
