"""Finite-difference operators of the explicit step (phasefield5d/solver/operators.py).

Device implementations with the reference signatures: periodic Laplacian,
serial-semantics abs-max (NaN ignored), in-place simplex clamp.
"""
import ctypes as C

import numpy as np

from .. import _lib
from ..device import Staged, prec_of


def calculate_laplacian(array, cell_size, out=None):
    """Second-order periodic Laplacian (operators.py:10-81); ``out`` written in place."""
    dim = array.ndim - 1
    if dim not in (1, 2, 3):
        raise NotImplementedError(f"Laplacian not implemented for spatial_dims={dim}")
    prec = prec_of(array)
    X = Staged(array)
    if out is None:
        out = np.empty_like(array) if X.host is not None else X.dev.new_empty(X.dev.shape)
    O = Staged(out, copy_in=False)
    shape = _lib.larr(X.dev.shape[:-1])
    Xd = X.dev.contiguous()
    _lib.check(_lib.lib().pf5d_laplacian(prec, C.c_void_p(Xd.data_ptr()), dim, shape, float(cell_size),
                                        O.ptr(), _lib.stream_handle()))
    return O.finish()


def abs_max(arr):
    """max |arr| over all elements, NaN ignored (operators.py:161-175)."""
    prec = prec_of(arr)
    A = Staged(arr)
    Ad = A.dev.contiguous()
    res = C.c_double()
    _lib.check(_lib.lib().pf5d_abs_max(prec, C.c_void_p(Ad.data_ptr()), Ad.numel(), C.byref(res),
                                      _lib.stream_handle()))
    return res.value


def clamp_simplex(composition):
    """Renormalise cells with sum(X) > 1, in place (operators.py:262-317)."""
    dim = composition.ndim - 1
    if dim not in (1, 2, 3):
        raise NotImplementedError(f"clamp_simplex not implemented for spatial_dims={dim}")
    prec = prec_of(composition)
    X = Staged(composition)
    if not X.dev.is_contiguous():
        raise ValueError("composition must be contiguous")
    _lib.check(_lib.lib().pf5d_clamp_simplex(prec, X.ptr(), X.dev.numel() // 4, _lib.stream_handle()))
    X.finish()
    return composition


def update_composition(composition, composition_change, dt):
    """composition_change *= dt; composition += composition_change (simulate.py:316-317)."""
    prec = prec_of(composition)
    X = Staged(composition)
    D = Staged(composition_change)
    _lib.check(_lib.lib().pf5d_update(prec, X.ptr(), D.ptr(), X.dev.numel(), float(dt), _lib.stream_handle()))
    D.finish()
    X.finish()
    return composition


def kappa_subtract(laplacian, mu, kappa):
    """laplacian *= kappa; mu -= laplacian (simulate.py:254-255), mu may be a strided view."""
    from ..device import row_stride
    prec = prec_of(laplacian)
    L = Staged(laplacian)
    M = Staged(mu)
    k = _lib.darr(np.asarray(kappa, dtype=float).reshape(-1)[:4])
    _lib.check(_lib.lib().pf5d_kappa_subtract(prec, L.ptr(), M.ptr(), row_stride(M.dev, 4),
                                             L.dev.numel() // 4, k, _lib.stream_handle()))
    L.finish()
    M.finish()
    return mu
