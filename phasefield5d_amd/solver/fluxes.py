"""Face fluxes (phasefield5d/solver/fluxes.py:6-125) and the fused gradient +
face flux + divergence (:296-416)."""
import ctypes as C

import numpy as np

from .. import _lib
from ..device import Staged, prec_of, row_stride


def _make(dim):
    def compute_divergence_from_mu(current_composition, mobilities, mu, inv_dx, out):
        if current_composition.ndim - 1 != dim:
            raise ValueError(f"expected a {dim}D field")
        prec = prec_of(current_composition)
        X = Staged(current_composition)
        B = Staged(mobilities)
        M = Staged(mu)
        O = Staged(out, copy_in=False)
        Xd = X.dev.contiguous()
        shape = _lib.larr(Xd.shape[:-1])
        _lib.check(_lib.lib().pf5d_divergence_from_mu(
            prec, C.c_void_p(Xd.data_ptr()), B.ptr(), row_stride(B.dev, 5), M.ptr(), row_stride(M.dev, 4),
            dim, shape, float(inv_dx), O.ptr(), _lib.stream_handle()))
        return O.finish()
    compute_divergence_from_mu.__name__ = f"_compute_divergence_from_mu_{dim}d"
    return compute_divergence_from_mu


_FUNCS = {d: _make(d) for d in (1, 2, 3)}


def get_divergence_from_mu_function(system_dim):
    """f(X, Mob, mu, inv_dx, out) writing dX/dt; Mob and mu may be row-strided views."""
    if system_dim not in _FUNCS:
        raise ValueError(f"Unsupported system_dim={system_dim}. Choose 1, 2, or 3.")
    return _FUNCS[system_dim]


def _make_flux(dim):
    def compute_fluxes_pm(current_composition, mobilities, grad_mu_plus, grad_mu_minus, fluxes_p, fluxes_m):
        if current_composition.ndim - 1 != dim:
            raise ValueError(f"expected a {dim}D field")
        prec = prec_of(current_composition)
        X = Staged(current_composition)
        B = Staged(mobilities)
        GP = Staged(grad_mu_plus)
        GM = Staged(grad_mu_minus)
        FP = Staged(fluxes_p, copy_in=False)
        FM = Staged(fluxes_m, copy_in=False)
        Xd = X.dev.contiguous()
        gp, gm = GP.dev.contiguous(), GM.dev.contiguous()
        shape = _lib.larr(Xd.shape[:-1])
        _lib.check(_lib.lib().pf5d_fluxes_pm(
            prec, C.c_void_p(Xd.data_ptr()), B.ptr(), row_stride(B.dev, 5), dim, shape,
            C.c_void_p(gp.data_ptr()), C.c_void_p(gm.data_ptr()), FP.ptr(), FM.ptr(), _lib.stream_handle()))
        FP.finish()
        FM.finish()
    compute_fluxes_pm.__name__ = f"_compute_fluxes_pm_{dim}d"
    return compute_fluxes_pm


_FLUX = {d: _make_flux(d) for d in (1, 2, 3)}


def get_flux_function(system_dim):
    """f(X, Mob, grad_mu_p, grad_mu_m, fluxes_p, fluxes_m) filling the (dim, *spatial, 4)
    flux buffers in place (fluxes.py:6-125); Mob may be a row-strided view."""
    if system_dim not in _FLUX:
        raise ValueError(f"Unsupported system_dim={system_dim}. Choose 1, 2, or 3.")
    return _FLUX[system_dim]
