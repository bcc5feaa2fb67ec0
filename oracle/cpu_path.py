"""The reference step loop on the CPU through the C/OpenMP restatement -- TEST INFRASTRUCTURE.

Used by bench.py's cpu_baseline leg (timing) and by tests (agreement with the
NumPy oracle).  Per step, in the order of examples/simulate.py:242-320:
interpolation (+ KD-tree rows for cells past the linear region), Laplacian,
L *= kappa, mu -= L, elastic (misfit -> scipy.fft rfftn/irfftn -> mu += psi l),
fused divergence, abs_max, adaptive dt, dX *= dt, X += dX, clamp.
"""
import ctypes as C
import os

import numpy as np
import scipy.fft as sfft

from oracle import pf5d_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle_pf5d.so")

_lib = None
pd = C.POINTER(C.c_double)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from oracle.build import build
            build()
        h = C.CDLL(LIB)
        h.ora_interp.restype = C.c_long
        h.ora_interp.argtypes = [pd, C.c_long, pd, C.c_int, C.c_int, C.c_double, pd]
        h.ora_laplacian.argtypes = [pd, C.c_int, C.POINTER(C.c_long), C.c_double, pd]
        h.ora_kappa_sub.argtypes = [pd, pd, C.c_long, pd]
        h.ora_misfit.argtypes = [pd, C.c_long, pd, pd, pd]
        h.ora_add_psi.argtypes = [pd, pd, C.c_long, pd, pd]
        h.ora_div_from_mu.argtypes = [pd, pd, C.c_int, C.POINTER(C.c_long), C.c_double, pd]
        h.ora_abs_max.restype = C.c_double
        h.ora_abs_max.argtypes = [pd, C.c_long]
        h.ora_update.argtypes = [pd, pd, C.c_long, C.c_double]
        h.ora_clamp.argtypes = [pd, C.c_long]
        h.ora_threads.restype = C.c_int
        h.ora_set_threads.argtypes = [C.c_int]
        _lib = h
    return _lib


def _p(a):
    return a.ctypes.data_as(pd)


class CPath:
    """Preallocated buffers + one reference step per call (FP64)."""

    def __init__(self, grid, resolution, tree, values, h, kappa, X0, l_vec=None, Bhat=None, ctld=1.0,
                 dt_cfl=np.inf, lower=1e-5, upper=1e-3, threads=None):
        self.L = lib()
        if threads:
            self.L.ora_set_threads(int(threads))
        self.grid = np.ascontiguousarray(grid, dtype=np.float64)
        self.num = grid.shape[0]
        self.res, self.tree, self.values = float(resolution), tree, values
        self.h = float(h)
        self.kappa = np.ascontiguousarray(np.asarray(kappa, dtype=np.float64).reshape(4))
        self.X0 = np.ascontiguousarray(X0, dtype=np.float64)
        self.l = None if l_vec is None else np.ascontiguousarray(l_vec, dtype=np.float64)
        self.Bhat = Bhat
        self.ctld, self.dt_cfl, self.lower, self.upper = ctld, dt_cfl, lower, upper
        self._shape = None

    def threads(self):
        return self.L.ora_threads()

    def _alloc(self, shape):
        n = int(np.prod(shape))
        self.P = np.empty((n, 9))
        self.Lap = np.empty((n, 4))
        self.dX = np.empty((n, 4))
        self.s = np.empty(n)
        self.tmp = np.empty((n, 4))
        self.shp = (C.c_long * 3)(*list(shape) + [1] * (3 - len(shape)))
        self._shape = shape

    def rate(self, X):
        shape = X.shape[:-1]
        if self._shape != shape:
            self._alloc(shape)
        n = int(np.prod(shape))
        Xf = np.ascontiguousarray(X).reshape(n, 4)
        nbad = self.L.ora_interp(_p(Xf), n, _p(self.grid), self.num, 9, self.res, _p(self.P))
        if nbad:
            bad = ~(((Xf[:, 0] + Xf[:, 1]) + Xf[:, 2]) + Xf[:, 3] <= O.linear_threshold(self.res))
            _, nn = self.tree.query(Xf[bad])
            self.P[bad] = self.values[nn]
        self.L.ora_laplacian(_p(Xf), len(shape), self.shp, self.h, _p(self.Lap))
        self.L.ora_kappa_sub(_p(self.Lap), _p(self.P), n, _p(self.kappa))
        if self.l is not None and self.Bhat is not None:
            self.L.ora_misfit(_p(Xf), n, _p(self.X0), _p(self.l), _p(self.s))
            axes = tuple(range(len(shape)))
            k = sfft.rfftn(self.s.reshape(shape), axes=axes, workers=1)
            k *= self.Bhat
            psi = np.ascontiguousarray(sfft.irfftn(k, s=shape, axes=axes, workers=1).real)
            self.L.ora_add_psi(_p(self.P), _p(psi.reshape(-1)), n, _p(self.l), _p(self.tmp))
        self.L.ora_div_from_mu(_p(Xf), _p(self.P), len(shape), self.shp, 1.0 / self.h, _p(self.dX))
        return self.dX.reshape(X.shape)

    def step(self, X, dt):
        """In place: X <- clamp(X + dt_new dX/dt); returns (dt_new, xmax)."""
        dX = self.rate(X)
        n = X.size
        xmax = self.L.ora_abs_max(_p(self.dX), n)
        dt_new = O.apply_adaptive_time(dt, xmax, self.ctld, self.dt_cfl, self.lower, self.upper)
        Xf = X.reshape(-1, 4)
        self.L.ora_update(_p(Xf), _p(self.dX), n, dt_new)
        self.L.ora_clamp(_p(Xf), n // 4)
        return dt_new, xmax
