"""Linear stability analysis of the multicomponent Cahn-Hilliard operator (host side)."""
from .operators import (  # noqa: F401
    calculate_wavenumber_cutoff,
    calculate_kinetic_operators,
    calculate_kinetic_eigenvalues_vectors,
)
