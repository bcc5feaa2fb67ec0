"""Scalar cubic elastic anisotropy of Khachaturyan type.

phasefield5d/elasticity/anisotropy.py:35-65.  B(n) is the orientation-dependent
part of the coherency energy kernel relative to its orientational average; it is
the scalar form of the Fourier-space kernel built by
``phasefield5d_amd.solver.elastic.build_khachaturyan_kernel``.
"""
import numpy as np


def _anisotropy_terms(c11, c12, c44):
    a = c11 - c12 - 2.0 * c44
    b = c11 + c12 + 2.0 * c44
    d = c11 + 2.0 * c12 + 4.0 * c44
    return a, b, d


def cubic_phi(n, c11, c12, c44):
    """phi(n) of the exact cubic Green's function for unit direction n."""
    n = np.asarray(n, dtype=float)
    n1, n2, n3 = n / np.linalg.norm(n)
    a, b, _ = _anisotropy_terms(c11, c12, c44)
    q = a / c44
    s2 = n1**2 * n2**2 + n2**2 * n3**2 + n3**2 * n1**2
    s3 = n1**2 * n2**2 * n3**2
    return c11 * (1 + 2 * q * s2 + 3 * q**2 * s3) / (c11 + a * (c11 + c12) / c44 * s2 + q**2 * b * s3)


def cubic_phi_average(c11, c12, c44):
    """Orientational average of phi(n)."""
    a, b, d = _anisotropy_terms(c11, c12, c44)
    return 1 + 4 * a / b / 5 + 54 * a**2 / (b * d) / 105


def calculate_khachaturyan_elastic_anisotropy(n, c11, c12, c44):
    """Return ``(B_n, Y_n)`` in GPa for direction ``n``."""
    phi = cubic_phi(n, c11, c12, c44)
    pre = (c11 + 2 * c12) ** 2 / c11
    return -pre * (phi - cubic_phi_average(c11, c12, c44)), 3 * (c11 + 2 * c12) - pre * phi
