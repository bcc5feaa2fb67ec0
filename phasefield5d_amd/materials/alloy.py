"""Alloy constant assembly (phasefield5d/materials/alloy.py:20-77)."""
import numpy as np

from .database import RADIUS_SOURCES, get_element_data


def build_alloy_constants(elements, radius_source="senkov"):
    """Return ``(alloy, ri, mu, nu, Vi, qi)`` for ``elements`` (first = dependent).

    ``qi = 2 mu (1 + nu) / (1 - nu)`` in GPa, as in the reference.
    """
    if radius_source not in RADIUS_SOURCES:
        raise ValueError(f"radius_source='{radius_source}' not recognised. "
                         f"Choose from {RADIUS_SOURCES}.")
    rows = [get_element_data(e) for e in elements]
    ri = np.array([r[f"atomic_radius_{radius_source}"] for r in rows], dtype=float)
    mu = np.array([r["shear_modulus_gpa"] for r in rows], dtype=float)
    nu = np.array([r["poissons_ratio"] for r in rows], dtype=float)
    vi = np.array([r["molar_volume_m3mol"] for r in rows], dtype=float)
    qi = 2.0 * mu * (1.0 + nu) / (1.0 - nu)
    return list(elements), ri, mu, nu, vi, qi
