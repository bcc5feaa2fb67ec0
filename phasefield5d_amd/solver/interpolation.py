"""4D CALPHAD interpolation (phasefield5d/solver/interpolation.py).

``build_4d_grid`` / ``build_calphad_kdtree`` are host setup (once per run) and
keep the reference signatures.  ``interpolator_nb`` keeps the reference
signature and runs the device kernel ``pf5d_interpolate`` (linear region and
nearest-node fallback both on the GPU).  ``interpolate_points_host`` serves
only the handful of setup evaluations at X0 and X0 +- h e_i
(examples/simulate.py:80-87, solver/system.py:284-293) and is never used on
fields.
"""
import ctypes as C

import numpy as np
from scipy.spatial import cKDTree

from .. import _lib
from ..device import Staged, device_table, is_device, prec_of, row_stride


def build_4d_grid(original_data, resolution, dtype=np.float64):
    """Dense ``(num,)*4 + (n_props,)`` grid with NaN outside the table (interpolation.py:13-35)."""
    data = original_data.to_numpy() if hasattr(original_data, "to_numpy") else np.asarray(original_data)
    step = float(resolution)
    num = int(round(1.0 / step)) + 1
    xdim = np.linspace(0.0, 1.0, num)
    coords = data[:, :4].astype(np.float64, copy=False)
    values = data[:, 4:].astype(dtype, copy=False)
    idx = np.clip(np.rint(coords / step).astype(np.int32), 0, num - 1)
    grid = np.full((num,) * 4 + (values.shape[1],), np.nan, dtype=dtype)
    grid[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3], :] = values
    return xdim, grid


def build_calphad_kdtree(original_data):
    """KD-tree over the table compositions and the row values (interpolation.py:38-44)."""
    data = original_data.to_numpy() if hasattr(original_data, "to_numpy") else np.asarray(original_data, dtype=float)
    return cKDTree(data[:, :4].astype(float)), data[:, 4:].astype(float)


def linear_threshold(resolution):
    return 1.0 - 4.0 * resolution - 1e-9


def interpolate_points_host(points, data_grid, resolution, tree, calphad_values):
    """Setup-only evaluation at a few composition points (same rules as the device kernel)."""
    P = np.atleast_2d(np.asarray(points, dtype=np.float64))
    num = data_grid.shape[0]
    inv = 1.0 / resolution
    out = np.empty((P.shape[0], data_grid.shape[-1]))
    for r, x in enumerate(P):
        if ((x[0] + x[1]) + x[2]) + x[3] > linear_threshold(resolution):
            _, nn = tree.query(x)
            out[r] = calphad_values[nn]
            continue
        ii, tt = [], []
        for d in range(4):
            u = x[d] * inv
            if u <= 0.0:
                i, t = 0, 0.0
            elif u >= num - 1:
                i, t = num - 2, 1.0
            else:
                i = int(u)
                t = u - i
            ii.append(i)
            tt.append(t)
        acc = None
        for b in range(16):
            bits = [(b >> (3 - d)) & 1 for d in range(4)]
            w = (tt[0] if bits[0] else 1.0 - tt[0]) * (tt[1] if bits[1] else 1.0 - tt[1])
            w = w * (tt[2] if bits[2] else 1.0 - tt[2])
            w = w * (tt[3] if bits[3] else 1.0 - tt[3])
            term = w * data_grid[ii[0] + bits[0], ii[1] + bits[1], ii[2] + bits[2], ii[3] + bits[3], :]
            acc = term if acc is None else acc + term
        out[r] = acc
    return out


def _check_nn_distance(X, resolution, tree, max_nn_dist):
    Xh = X.detach().cpu().numpy() if is_device(X) else np.asarray(X)
    Xf = Xh.reshape(-1, 4).astype(np.float64)
    s = ((Xf[:, 0] + Xf[:, 1]) + Xf[:, 2]) + Xf[:, 3]
    bad = Xf[~(s <= linear_threshold(resolution))]
    if len(bad):
        d, _ = tree.query(bad)
        if np.any(d > max_nn_dist):
            raise RuntimeError("Some compositions are outside the CALPHAD domain (KD-tree).")


def interpolator_nb(composition_array, data_grid, resolution, tree, calphad_values, max_nn_dist=None,
                    out=None, mask_buffer=None):
    """mu(X), Mob(X) for every cell, written into ``out`` (…, 9) (interpolation.py:154-211).

    Runs on the GPU.  ``tree``/``calphad_values`` are only consulted for the
    optional ``max_nn_dist`` check; the nearest-node values come from the device
    grid.  ``mask_buffer`` is accepted for signature compatibility.
    """
    prec = prec_of(composition_array)
    X = Staged(composition_array, np.float64 if prec == _lib.F64 else np.float32)
    spatial = tuple(X.dev.shape[:-1])
    if X.dev.shape[-1] != 4:
        raise ValueError("composition_array must end in 4 components")
    if max_nn_dist is not None:
        _check_nn_distance(X.dev, resolution, tree, max_nn_dist)
    ncell = int(np.prod(spatial)) if spatial else 1
    if out is None:
        out = np.empty(spatial + (data_grid.shape[-1],), dtype=np.float64 if prec == _lib.F64 else np.float32) \
            if X.host is not None else X.dev.new_empty(spatial + (data_grid.shape[-1],))
    O = Staged(out, copy_in=False)
    tab = device_table(data_grid, resolution, prec)
    Xd = X.dev.contiguous()
    _lib.check(_lib.lib().pf5d_interpolate(prec, C.c_void_p(Xd.data_ptr()), ncell, tab.handle,
                                          O.ptr(), row_stride(O.dev, 9), _lib.stream_handle()))
    return O.finish()
