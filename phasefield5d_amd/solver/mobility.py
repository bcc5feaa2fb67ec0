"""Dyadic face-mobility matrices (phasefield5d/solver/mobility.py:36-78).

Not part of the step: the reference keeps these for external callers and its
own tests, and the step's flux kernels evaluate the same matrix-vector product
inline (``face_flux``).  Host NumPy, for inspection of small fields.
"""
import numpy as np


def _face_means(a):
    """Per axis d: the mean of each cell with its +e_d neighbour (forward face) and with
    its -e_d neighbour (backward face), periodic; (dim, *spatial, n) each."""
    dim = a.ndim - 1
    fwd = np.stack([0.5 * (a + np.roll(a, -1, axis=d)) for d in range(dim)])
    bwd = np.stack([np.roll(fwd[d], 1, axis=d) for d in range(dim)])
    return fwd, bwd


def _dyadic(Xm, Mm):
    """M = w_Fe v v^T + sum_k w_k (e_k - X)(e_k - X)^T with v = -X, w_Fe = (1 - sum X) Mob_Fe,
    w_k = X_k Mob_k, at face compositions Xm (..., S) and mobilities Mm (..., S+1)."""
    S = Xm.shape[-1]
    w_fe = (1.0 - Xm.sum(axis=-1)) * Mm[..., 0]
    out = w_fe[..., None, None] * (Xm[..., :, None] * Xm[..., None, :])
    for k in range(S):
        u = -Xm.copy()
        u[..., k] += 1.0
        out = out + (Xm[..., k] * Mm[..., 1 + k])[..., None, None] * (u[..., :, None] * u[..., None, :])
    return out


def calculate_mobility_matrix_pm(current_composition, mobilities):
    """(mobility_plus, mobility_minus), each (dim, *spatial, S, S): the dyadic mobility at
    the forward and backward faces of every cell along every axis.  ``mobilities`` is
    ordered [Fe, solute0, ..., solute(S-1)]."""
    Xf, Xb = _face_means(np.asarray(current_composition, dtype=np.float64))
    Mf, Mb = _face_means(np.asarray(mobilities, dtype=np.float64))
    return _dyadic(Xf, Mf), _dyadic(Xb, Mb)
