"""Save-step diagnostics of examples/simulate.py:296-346, computed on the device.

Before the update (from the step's own state): ``net`` = sum of dX/dt, the
per-component max |dX/dt|, and ``total_flux`` = sum of fluxes_p - fluxes_m from
the forward/backward gradients of mu and the face fluxes (calculate_gradients_pm
+ get_flux_function).  After the update: mass, per-component min / max / variance,
the minimum Fe fraction 1 - max sum(X), and the gradient-energy proxy
f_grad = -(h^dim / 2) sum X . kappa lap X.  All reductions accumulate in FP64.
"""
import numpy as np

from .device import torch
from .solver.fluxes import get_flux_function
from .solver.operators import calculate_gradients_pm, calculate_laplacian, field_dot, field_stats


def pre_update(X, mu, mob, dXdt, h):
    """Diagnostics of one step from its state before the update (device tensors)."""
    dim = X.dim() - 1
    d = field_stats(dXdt)
    gp, gm = calculate_gradients_pm(mu, h)
    fp, fm = torch().empty_like(gp), torch().empty_like(gm)
    get_flux_function(dim)(X, mob, gp, gm, fp, fm)
    tot = field_stats(fp.reshape(-1, 4))["sum"] - field_stats(fm.reshape(-1, 4))["sum"]
    return {"net": d["sum"], "max_per_comp": np.maximum(np.abs(d["min"]), np.abs(d["max"])), "total_flux": tot}


def post_update(X, h, kappa):
    """Diagnostics of the state after the update (device tensor)."""
    dim = X.dim() - 1
    s = field_stats(X)
    L = calculate_laplacian(X, h)
    f_grad = -0.5 * h ** dim * float(np.dot(field_dot(X, L), np.asarray(kappa, dtype=float).reshape(-1)))
    return {"mass": s["sum"], "comp_min": s["min"], "comp_max": s["max"], "comp_var": s["var"],
            "comp_fe_min": 1.0 - float(s["max_row_sum"]), "f_grad": f_grad}
