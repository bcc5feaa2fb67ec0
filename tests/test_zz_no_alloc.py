"""pf5d_bounds_stats_noalloc (fixed device scratch, no allocation) against pf5d_bounds_stats.

Both run the same partials kernel with the same grid and fold the partials in
the same order on the host, so the five statistics must agree bit for bit.  Two
calls in a row, the second on a smaller field, check that the shared scratch
left by the first call does not leak into the second.
"""
import ctypes as C

import numpy as np
import pytest

from phasefield5d_amd import _lib


def _field(ncell, seed, dtype):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-0.01, 0.4, size=(ncell, 4))
    X[ncell // 3, 2] = np.nan          # a non-finite cell
    X[ncell // 2] = (0.5, 0.4, 0.2, 0.1)  # sum above 1 + tol
    X[ncell // 5, 1] = -0.02           # a negative component
    return X.astype(dtype)


def _stats(fn, prec, Xd, ncell, stream):
    out = (C.c_double * 5)()
    _lib.check(fn(prec, C.c_void_p(Xd.data_ptr()), ncell, out, stream))
    return np.array(out[:], dtype=np.float64)


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_bounds_stats_equals_allocating(prec):
    import torch
    _lib.require_gpu()
    L = _lib.lib()
    dtype = np.float64 if prec == "fp64" else np.float32
    p = _lib.F64 if prec == "fp64" else _lib.F32
    st = _lib.stream_handle()
    for ncell, seed in ((300_001, 1), (4_097, 2), (1, 3)):
        print(f"[no-alloc] {prec} ncell={ncell}", flush=True)
        Xh = _field(ncell, seed, dtype) if ncell > 10 else np.full((1, 4), 0.2, dtype=dtype)
        Xd = torch.from_numpy(Xh).to("cuda")
        a = _stats(L.pf5d_bounds_stats_noalloc, p, Xd, ncell, st)
        b = _stats(L.pf5d_bounds_stats, p, Xd, ncell, st)
        assert np.array_equal(a, b), (ncell, a, b)
        # the host's own reading of the same field
        Xf = Xh.astype(np.float64)
        fin = np.isfinite(Xf).all(axis=1)
        assert a[0] == float((~fin).sum())
        s = ((Xf[:, 0] + Xf[:, 1]) + Xf[:, 2]) + Xf[:, 3]
        assert a[1] == s[fin].max() and a[2] == s[fin].min()
        assert a[3] == np.nanmin(Xf)
        assert a[4] == np.nanmax(Xf)
