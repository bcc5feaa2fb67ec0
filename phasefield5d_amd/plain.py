"""The reference loop body, one drop-in device call per reference function.

This is the "plain" FP64 GPU path: examples/simulate.py:242-320 with every
Numba/NumPy step replaced by the libpf5d entry point of the same name, the
fields kept on the device, and the host computing dt from abs_max exactly as
the reference does (one device->host scalar per step).
"""
import numpy as np

from .device import torch
from .solver.elastic import make_elastic_updater
from .solver.fluxes import get_divergence_from_mu_function
from .solver.interpolation import interpolator_nb
from .solver.operators import abs_max, calculate_laplacian, clamp_simplex, kappa_subtract, update_composition
from .solver.system import apply_adaptive_time, out_of_bounds


class PlainStepper:
    def __init__(self, pb, dtype=np.float64):
        t = torch()
        self.pb = pb
        cfg = pb.cfg
        tdt = t.float64 if dtype == np.float64 else t.float32
        self.X = t.from_numpy(np.ascontiguousarray(pb.X_init, dtype=dtype)).cuda()
        spatial = tuple(pb.shape)
        self.P = t.empty(spatial + (9,), dtype=tdt, device="cuda")
        self.L = t.empty_like(self.X)
        self.dX = t.empty_like(self.X)
        self.elastic_update = make_elastic_updater(cfg, pb.Lambda, pb.Bhat, self.X.shape, dtype)
        self.div = get_divergence_from_mu_function(len(spatial))
        self.kappa = pb.kappa.reshape(-1)
        self.dt = pb.dt0
        self.t = 0.0
        self.timestep = 0
        self.xmax = 0.0
        self.stopped = False

    def potentials(self):
        """mu and Mob of the current state into P (simulate.py:242-263); returns (mu, Mob) views."""
        pb, cfg = self.pb, self.pb.cfg
        interpolator_nb(self.X, pb.data_grid, cfg.resolution, pb.tree, pb.calphad_values, out=self.P)
        mu = self.P[..., :4]
        mob = self.P[..., 4:]
        calculate_laplacian(self.X, pb.dx, self.L)
        kappa_subtract(self.L, mu, self.kappa)
        self.elastic_update(mu, self.X)
        return mu, mob

    def step(self, oob_interval=100, save=None):
        """One reference step; returns False (and stops) on a hard bounds violation.

        ``save(timestep, t_after) -> (save_data, save_snapshot)`` is asked once dt is
        known (simulate.py:290-297); on a data save the step's diagnostics are left in
        ``self.last_diag`` (pre-update ones from this step's state, post-update ones after
        the clamp, as simulate.py:299-346 computes them)."""
        if self.stopped:
            return False
        pb, cfg = self.pb, self.pb.cfg
        if oob_interval and self.timestep % oob_interval == 0 and out_of_bounds(self.X):
            self.stopped = True
            return False
        mu, mob = self.potentials()
        self.div(self.X, mob, mu, 1.0 / pb.dx, self.dX)
        self.xmax = float(abs_max(self.dX))
        self.dt = apply_adaptive_time(self.dt, self.xmax, pb.ctld, pb.dt_cfl,
                                      composition_change_lower_limit=cfg.lower_limit,
                                      composition_change_upper_limit=cfg.upper_limit)
        self.last_diag = None
        sd = False
        if save is not None:
            sd, _ = save(self.timestep, self.t + self.dt)
        if sd:
            from . import diagnostics
            self.last_diag = diagnostics.pre_update(self.X, mu, mob, self.dX, pb.dx)
        self.t += self.dt
        update_composition(self.X, self.dX, self.dt)
        clamp_simplex(self.X)
        if sd:
            self.last_diag.update(diagnostics.post_update(self.X, pb.dx, self.kappa))
        self.timestep += 1
        return True
