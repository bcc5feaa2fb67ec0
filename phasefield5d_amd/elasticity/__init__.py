"""Vegard misfit, molar volume and cubic (Khachaturyan) anisotropy helpers."""
from .constants import load_material_constants, load_elastic_constants  # noqa: F401
from .energies import calculate_lattice_strain_coefficients, calculate_molar_volume  # noqa: F401
from .anisotropy import calculate_khachaturyan_elastic_anisotropy  # noqa: F401
