"""Vegard-law strain coefficients and mixed molar volume.

phasefield5d/elasticity/energies.py:5-26.
"""
import numpy as np


def calculate_lattice_strain_coefficients(X, ri):
    """l_i = (r_i - r_Fe) / r_mix for the four independent components."""
    X = np.asarray(X, dtype=float)
    x_dep = 1.0 - np.sum(X)
    r_mix = ri[0] * x_dep + np.dot(ri[1:], X)
    return (ri[1:] - ri[0]) / r_mix


def calculate_molar_volume(X, Vi):
    """Linear mixing of the five molar volumes [m^3/mol]."""
    X = np.asarray(X, dtype=float)
    x_dep = 1.0 - np.sum(X)
    return Vi[0] * x_dep + np.dot(Vi[1:], X)
