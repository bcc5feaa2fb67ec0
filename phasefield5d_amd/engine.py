"""Device-resident fused step over one or many simulations (libpf5d engine).

One ``Engine`` holds any number of simulations (slots) of any dimension and
size.  A step enqueues, as one captured HIP graph: the FFT passes of every
elastic simulation, one fused stencil launch (K1) over the tiles of all
simulations, one dt-controller launch (one block per simulation) and one
update launch (K2).  dt, time, max|dX/dt|, FP64 mass and bounds statistics stay
on the device; the host reads them only when asked.
"""
import ctypes as C
import math

import numpy as np

from . import _lib
from .device import device_table

STATE_FIELDS = ("dt", "t", "last_xmax", "mass0", "mass1", "mass2", "mass3", "steps", "status", "sum_max",
                "sum_min", "comp_min", "comp_max", "nonfinite", "mean0", "mean1")
ST_ACTIVE, ST_HARD_OOB, ST_SOFT_OOB, ST_DONE = 1, 2, 4, 8


class Engine:
    def __init__(self, precision="fp64", stream=None):
        _lib.require_gpu()
        self.prec = _lib.F32 if precision in ("fp32", _lib.F32) else _lib.F64
        self.precision = "fp32" if self.prec == _lib.F32 else "fp64"
        h = C.c_void_p()
        _lib.check(_lib.lib().pf5d_engine_create(self.prec, stream, C.byref(h)))
        self.h = h
        self.slots = {}

    def close(self):
        h = getattr(self, "h", None)
        self.h = None
        if h is not None and h.value and _lib is not None and _lib._lib is not None:
            _lib._lib.pf5d_engine_destroy(h)

    def __del__(self):
        self.close()

    # -- membership ------------------------------------------------------------
    def add_problem(self, pb, X=None, t_end=math.inf, oob_interval=100, dt0=None, tol_soft=1e-5, tol_hard=1e-3):
        """Add a set-up problem (see problem.build_problem); returns its slot.  A problem
        without a host field (X_init None) and with ``extra["device_init"] = (seed,
        fluctuation)`` has its initial field drawn on the device.  Every ``oob_interval``
        steps the field is checked with the tolerances of ``out_of_bounds``
        (system.py:12-53): past ``tol_hard`` the simulation stops."""
        cfg = pb.cfg
        d = _lib.SimDesc()
        d.dim = len(pb.shape)
        for a, n in enumerate(pb.shape):
            d.shape[a] = n
        d.cell_size = pb.dx
        for a in range(4):
            d.kappa[a] = float(pb.kappa.reshape(-1)[a])
            d.X0[a] = float(pb.X0[a])
            d.l[a] = float(pb.l_vec[a]) if pb.l_vec is not None else 0.0
        d.elastic = 1 if pb.Bhat is not None else 0
        d.ctld, d.dt_cfl = pb.ctld, pb.dt_cfl
        d.dt0 = pb.dt0 if dt0 is None else dt0
        d.lower_limit, d.upper_limit = cfg.lower_limit, cfg.upper_limit
        d.t_end = t_end if math.isfinite(t_end) else 1e300
        d.oob_interval = oob_interval
        d.tol_soft, d.tol_hard = float(tol_soft), float(tol_hard)
        bh = np.ascontiguousarray(pb.Bhat, dtype=np.float64) if pb.Bhat is not None else None
        tab = device_table(pb.data_grid, cfg.resolution, self.prec)
        slot = C.c_int()
        gen = pb.extra.get("device_init") if X is None and pb.X_init is None else None
        if gen is not None:
            # the field is drawn on the device: (seed, fluctuation)
            seed, fluct = gen
            _lib.check(_lib.lib().pf5d_engine_add_generated(self.h, C.byref(d), int(seed), float(fluct),
                                                           _lib.ptr_f64(bh) if bh is not None else None,
                                                           tab.handle, C.byref(slot)))
        else:
            Xh = np.ascontiguousarray(pb.X_init if X is None else X, dtype=np.float64)
            if Xh.shape != tuple(pb.shape) + (4,):
                raise ValueError(f"X shape {Xh.shape} != {tuple(pb.shape) + (4,)}")
            _lib.check(_lib.lib().pf5d_engine_add(self.h, C.byref(d), _lib.ptr_f64(Xh),
                                                 _lib.ptr_f64(bh) if bh is not None else None, tab.handle,
                                                 C.byref(slot)))
        self.slots[slot.value] = (pb, tab)
        return slot.value

    def remove(self, slot):
        _lib.check(_lib.lib().pf5d_engine_remove(self.h, slot))
        self.slots.pop(slot, None)

    # -- stepping ----------------------------------------------------------------
    def step(self, n=1):
        _lib.check(_lib.lib().pf5d_engine_step(self.h, int(n)))

    def prepare(self, n=1):
        """Capture the step graphs ``step(n)`` replays, without running them."""
        _lib.check(_lib.lib().pf5d_engine_prepare(self.h, int(n)))

    def sync(self):
        _lib.check(_lib.lib().pf5d_engine_sync(self.h))

    def variants(self):
        n = C.c_int()
        _lib.check(_lib.lib().pf5d_engine_num_variants(self.h, C.byref(n), None))
        info = (C.c_int * (5 * n.value))()
        _lib.check(_lib.lib().pf5d_engine_num_variants(self.h, C.byref(n), info))
        return [tuple(info[5 * i:5 * i + 5]) for i in range(n.value)]

    def set_tiling(self, variant, cx):
        _lib.check(_lib.lib().pf5d_engine_set_tiling(self.h, int(variant), int(cx)))

    def set_fft(self, budget_kb, nt, parts=None, plane_nt=256, wave=None):
        """Line-pass batching (LDS KiB per item, threads per block) and, when given,
        the parts per plane of the 3D plane passes (0 = line passes only), their
        block size, and which passes run wave-resident where lengths are 128 (mask:
        1 forward plane pass, 2 x pass, 4 inverse plane pass; pf5d_engine_set_fft_wave)."""
        _lib.check(_lib.lib().pf5d_engine_set_fft(self.h, int(budget_kb), int(nt)))
        if parts is not None:
            _lib.check(_lib.lib().pf5d_engine_set_fft_planes(self.h, int(parts), int(plane_nt)))
        if wave is not None:
            _lib.check(_lib.lib().pf5d_engine_set_fft_wave(self.h, int(wave)))

    # -- readback ----------------------------------------------------------------
    def state(self, slot):
        buf = (C.c_double * 16)()
        _lib.check(_lib.lib().pf5d_engine_get_state(self.h, slot, buf))
        s = dict(zip(STATE_FIELDS, list(buf)))
        s["mass"] = np.array([s.pop(f"mass{i}") for i in range(4)])
        s["steps"] = int(s["steps"])
        s["status"] = int(s["status"])
        return s

    def all_states(self, maxn=4096):
        """{slot: state dict} of every occupied slot, one device->host copy."""
        buf = (C.c_double * (16 * maxn))()
        n = C.c_int()
        _lib.check(_lib.lib().pf5d_engine_get_all_states(self.h, maxn, buf, C.byref(n)))
        out = {}
        for i in range(n.value):
            if i not in self.slots:
                continue
            v = list(buf[16 * i:16 * i + 16])
            s = dict(zip(STATE_FIELDS, v))
            s["mass"] = np.array([s.pop(f"mass{k}") for k in range(4)])
            s["steps"] = int(s["steps"])
            s["status"] = int(s["status"])
            out[i] = s
        return out

    def set_time(self, slot, dt, t):
        _lib.check(_lib.lib().pf5d_engine_set_time(self.h, slot, float(dt), float(t)))

    def field(self, slot, which="X"):
        pb = self.slots[slot][0]
        w = {"X": 0, "dXdt": 1, "psi": 2}[which]
        shape = tuple(pb.shape) + ((4,) if w < 2 else ())
        out = np.empty(shape, dtype=np.float64)
        _lib.check(_lib.lib().pf5d_engine_get_field(self.h, slot, w, _lib.ptr_f64(out)))
        return out

    def set_field(self, slot, X):
        Xh = np.ascontiguousarray(X, dtype=np.float64)
        _lib.check(_lib.lib().pf5d_engine_set_field(self.h, slot, _lib.ptr_f64(Xh)))

    def field_ptr(self, slot, which="X"):
        return _lib.lib().pf5d_engine_field_ptr(self.h, slot, {"X": 0, "dXdt": 1, "psi": 2}[which])

    def stream_ptr(self):
        return _lib.lib().pf5d_engine_stream(self.h)

    def stats(self):
        a, b = C.c_long(), C.c_long()
        _lib.check(_lib.lib().pf5d_engine_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def time_kernel(self, which="k1", reps=20):
        """Average milliseconds of one launch of a stage (HIP events on the engine stream)."""
        ms = C.c_double()
        w = {"k1": 0, "fft": 1, "dt": 2, "k2prime": 3, "k1_potentials": 4, "k1_flux": 5, "step_like": 6}[which]
        _lib.check(_lib.lib().pf5d_engine_time_kernel(self.h, w, int(reps), C.byref(ms)))
        return ms.value

    def time_step_stages(self, reps=50, separate_dt=False):
        """Average ms per stage over ``reps`` whole steps launched one by one with HIP
        events between the stages (one-group engine, marching two-pass stencil).  The
        steps advance the simulations as ``step(reps)`` does."""
        out = (C.c_double * 6)()
        _lib.check(_lib.lib().pf5d_engine_time_step_stages(self.h, int(reps), 1 if separate_dt else 0, out))
        return dict(zip(("fft", "k1_potentials", "k1_flux", "dt", "k2", "step"), list(out)))

    def num_groups(self):
        """Simulation groups stepped one after another (1: all step together)."""
        n = C.c_int()
        _lib.check(_lib.lib().pf5d_engine_num_groups(self.h, C.byref(n)))
        return n.value

    def launch_info(self):
        out = (C.c_long * 13)()
        _lib.check(_lib.lib().pf5d_engine_launch_info(self.h, out))
        return dict(zip(("k1_items", "k2_items", "fft_items", "k1_grid", "k1_lds", "variant", "cx", "fft_budget",
                         "fft_nt", "kf_grid", "fft_parts", "fft_plane_nt", "fft_wave"), list(out)))
