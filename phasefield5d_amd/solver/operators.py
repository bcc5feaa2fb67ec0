"""Finite-difference operators of the explicit step (phasefield5d/solver/operators.py).

Device implementations with the reference signatures: periodic Laplacian,
forward/backward gradients, serial-semantics abs-max (NaN ignored), in-place
simplex clamp, and the FP64 field statistics of the save steps.
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


def calculate_gradients_pm(array, cell_size, grad_p=None, grad_m=None):
    """Forward (+) and backward (-) differences along each axis (operators.py:88-158).

    ``array`` is (*spatial, 4), possibly a row-strided view (the mu columns of the
    interpolated data); returns (grad_p, grad_m) of shape (dim, *spatial, 4).
    """
    from ..device import row_stride
    dim = array.ndim - 1
    if dim not in (1, 2, 3):
        raise NotImplementedError(f"Gradients not implemented for spatial_dims={dim}")
    prec = prec_of(array)
    A = Staged(array)
    gshape = (dim,) + tuple(A.dev.shape)
    if grad_p is None:
        grad_p = np.empty(gshape, dtype=np.asarray(array[:0]).dtype) if A.host is not None else A.dev.new_empty(gshape)
    if grad_m is None:
        grad_m = np.empty_like(grad_p) if A.host is not None else A.dev.new_empty(gshape)
    P = Staged(grad_p, copy_in=False)
    M = Staged(grad_m, copy_in=False)
    shape = _lib.larr(A.dev.shape[:-1])
    _lib.check(_lib.lib().pf5d_gradients_pm(prec, A.ptr(), row_stride(A.dev, 4), dim, shape, float(cell_size),
                                           P.ptr(), M.ptr(), _lib.stream_handle()))
    return P.finish(), M.finish()


def field_stats(array):
    """FP64 statistics over the spatial axes of a (*spatial, n) field (n <= 8), for the
    save-step diagnostics of examples/simulate.py:299-346: dict of per-component
    ``sum``, ``min``, ``max``, ``var`` (population, about the mean) and ``max_row_sum``."""
    from ..device import row_stride
    prec = prec_of(array)
    A = Staged(array)
    n = A.dev.shape[-1]
    ncell = A.dev.numel() // n if A.dev.is_contiguous() else int(np.prod(A.dev.shape[:-1]))
    out = (C.c_double * (4 * n + 1))()
    _lib.check(_lib.lib().pf5d_field_stats(prec, A.ptr(), ncell, row_stride(A.dev, n), n, out,
                                          _lib.stream_handle()))
    v = np.frombuffer(out, dtype=np.float64).copy()
    return {"sum": v[:n], "min": v[n:2 * n], "max": v[2 * n:3 * n], "var": v[3 * n:4 * n], "max_row_sum": v[4 * n]}


def field_dot(a, b):
    """Per-component FP64 sums of a * b over the spatial axes of two (*spatial, n) fields."""
    prec = prec_of(a)
    A, B = Staged(a), Staged(b)
    Ad, Bd = A.dev.contiguous(), B.dev.contiguous()
    n = Ad.shape[-1]
    out = (C.c_double * n)()
    _lib.check(_lib.lib().pf5d_field_dot(prec, C.c_void_p(Ad.data_ptr()), C.c_void_p(Bd.data_ptr()),
                                        Ad.numel() // n, n, out, _lib.stream_handle()))
    return np.frombuffer(out, dtype=np.float64).copy()


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
