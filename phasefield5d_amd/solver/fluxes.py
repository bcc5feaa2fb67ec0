"""Fused gradient + face flux + divergence (phasefield5d/solver/fluxes.py:296-416)."""
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
