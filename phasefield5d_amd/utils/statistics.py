"""Significant-digit rounding used to derive the cell size and the initial time step.

Restates phasefield5d/utils/statistics.py:37-50 (round_to_first_nonzero,
round_down_to_first_nonzero).  Both feed host-side setup only (dx, dt0, L).
"""
import numpy as np


def _leading_exponent(x):
    return int(np.floor(np.log10(abs(x))))


def round_to_first_nonzero(x):
    """Round ``x`` to one significant digit (Python ``round`` semantics)."""
    if x == 0:
        return 0.0
    return round(x, -_leading_exponent(x))


def round_down_to_first_nonzero(x):
    """Truncate ``|x|`` to its first significant digit, keeping the sign."""
    if x == 0:
        return 0.0
    scale = 10 ** _leading_exponent(x)
    return np.sign(x) * np.floor(abs(x) / scale) * scale
