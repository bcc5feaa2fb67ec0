"""Device-array plumbing between the reference-shaped Python API and libpf5d.

Arrays handed to the drop-in functions may be NumPy arrays (copied to the GPU
and back, as a drop-in for the reference's in-place Numba calls) or torch
tensors already on the GPU (zero copy).  torch only provides device memory and
the stream; all arithmetic of the step runs in libpf5d's HIP kernels.
"""
import ctypes as C
import weakref

import numpy as np

from . import _lib

_TORCH = None


def torch():
    global _TORCH
    if _TORCH is None:
        import torch as _t
        _TORCH = _t
    return _TORCH


def is_device(a):
    t = torch()
    return isinstance(a, t.Tensor) and a.is_cuda


def prec_of(a):
    t = torch()
    if is_device(a):
        if a.dtype == t.float64:
            return _lib.F64
        if a.dtype == t.float32:
            return _lib.F32
        raise TypeError(f"unsupported device dtype {a.dtype}")
    return _lib.F32 if np.asarray(a).dtype == np.float32 else _lib.F64


class Staged:
    """A device view of an input array; ``finish()`` copies results back into NumPy outputs."""

    def __init__(self, a, dtype=None, copy_in=True):
        self.host = None
        if is_device(a):
            self.dev = a
        else:
            _lib.require_gpu()
            arr = np.asarray(a)
            if dtype is not None:
                arr = arr.astype(dtype, copy=False)
            self.host = a
            t = torch()
            if copy_in:
                self.dev = t.from_numpy(np.ascontiguousarray(arr)).to("cuda")
            else:
                self.dev = t.empty(arr.shape, dtype=t.from_numpy(arr[:0].copy()).dtype, device="cuda")

    def ptr(self):
        return C.c_void_p(self.dev.data_ptr())

    def finish(self):
        if self.host is not None and isinstance(self.host, np.ndarray):
            self.host[...] = self.dev.cpu().numpy()
        return self.host if self.host is not None else self.dev


def row_stride(view, width):
    """Row stride of a (…, width) view whose rows have unit element stride."""
    if view.stride(-1) != 1:
        raise ValueError("component axis must have unit stride")
    if view.dim() == 1:
        return width
    st = view.stride(-2)
    # the spatial axes must flatten with a uniform row stride (C-order view of a (N, rs) buffer)
    shape = view.shape[:-1]
    expect = st
    for ax in range(view.dim() - 2, -1, -1):
        if shape[ax] > 1 and view.stride(ax) != expect:
            raise ValueError("rows of the view are not evenly spaced")
        expect *= shape[ax]
    return st


class DeviceTable:
    """Padded device copy of the dense (num,num,num,num,9) property grid."""

    def __init__(self, data_grid, resolution, prec=_lib.F64):
        _lib.require_gpu()
        g = np.ascontiguousarray(np.asarray(data_grid, dtype=np.float64))
        if g.ndim != 5 or g.shape[0] != g.shape[1] or g.shape[-1] != 9:
            raise ValueError(f"data_grid must be (num,num,num,num,9), got {g.shape}")
        self.num = g.shape[0]
        self.resolution = float(resolution)
        self.prec = prec
        h = C.c_void_p()
        _lib.check(_lib.lib().pf5d_table_create(prec, self.num, 9, _lib.ptr_f64(g), self.resolution, C.byref(h)))
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib.lib().pf5d_table_destroy(h)
            self.handle = None


_TABLE_CACHE = {}


def device_table(data_grid, resolution, prec=_lib.F64):
    """Device table for a host grid, cached for the lifetime of the grid object."""
    key = (id(data_grid), float(resolution), prec)
    hit = _TABLE_CACHE.get(key)
    if hit is not None and hit[0]() is data_grid:
        return hit[1]
    tab = DeviceTable(data_grid, resolution, prec)
    try:
        ref = weakref.ref(data_grid, lambda _r, k=key: _TABLE_CACHE.pop(k, None))
    except TypeError:
        ref = (lambda g=data_grid: g)
    _TABLE_CACHE[key] = (ref, tab)
    return tab
