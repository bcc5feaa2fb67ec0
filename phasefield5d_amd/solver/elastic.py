"""Rank-1 Khachaturyan elastic driving force (phasefield5d/solver/elastic.py).

Host setup (kernel B(k), coupling matrix, rank-1 detection) follows
elastic.py:55-107 and :183-197.  The per-step update returned by
``make_elastic_updater`` runs on the GPU: misfit s = l.(X - X0), a hand-written
LDS Stockham FFT pair with the B(k) product fused into the last forward /
first inverse pass, and mu += psi l.
"""
import ctypes as C

import numpy as np

from .. import _lib
from ..device import Staged, prec_of, row_stride
from ..elasticity.energies import calculate_lattice_strain_coefficients, calculate_molar_volume
from ..elasticity.anisotropy import calculate_khachaturyan_elastic_anisotropy


def calculate_linear_elastic_coupling_matrix(reference_composition, ri):
    """Lambda = l l^T of the Vegard misfit strains (elastic.py:55-58)."""
    lih = calculate_lattice_strain_coefficients(reference_composition, ri)
    return np.outer(lih, lih)


def get_elastic_matrix(composition, vi, ri, c11, c12, c44, direction=np.array([1, 0, 0]), molar_volume=None):
    """B(direction) V_m Lambda in J/mol for the kinetics Hessian (elastic.py:61-66)."""
    vm = calculate_molar_volume(composition, vi) if molar_volume is None else molar_volume
    B_n, _ = calculate_khachaturyan_elastic_anisotropy(direction, c11, c12, c44)
    return calculate_linear_elastic_coupling_matrix(composition, ri) * B_n * 1e9 * vm


def build_khachaturyan_kernel(k_grid, k_norm, c11, c12, c44):
    """B(k) [J/m^3] of the exact cubic Green's function, B(0) = 0 (elastic.py:73-107)."""
    nz = k_norm > 0.0
    with np.errstate(invalid="ignore", divide="ignore"):
        n = [np.where(nz, k_grid[i] / k_norm, 0.0) for i in range(3)]
    a = c11 - c12 - 2.0 * c44
    b = c11 + c12 + 2.0 * c44
    d = c11 + 2.0 * c12 + 4.0 * c44
    q2 = [ni**2 for ni in n]
    s2 = q2[0] * q2[1] + q2[1] * q2[2] + q2[2] * q2[0]
    s3 = q2[0] * q2[1] * q2[2]
    phi = np.where(nz, c11 * (1.0 + 2.0 * a / c44 * s2 + 3.0 * (a / c44) ** 2 * s3)
                   / (c11 + a * (c11 + c12) / c44 * s2 + (a / c44) ** 2 * b * s3), 1.0)
    phi_avg = 1.0 + 4.0 * a / b / 5.0 + 54.0 * a**2 / (b * d) / 105.0
    return np.where(nz, -(c11 + 2.0 * c12) ** 2 / c11 * (phi - phi_avg), 0.0) * 1e9


def rank_one_vector(linear_elastic_coupling_matrix):
    """(l, is_rank1): l = v_max sqrt(lambda_max) when Lambda has one eigenvalue above 1e-10 max."""
    w, v = np.linalg.eigh(linear_elastic_coupling_matrix)
    largest = w[-1]
    n_nonzero = int(np.sum(w > 1e-10 * max(largest, 1e-30)))
    if n_nonzero == 1 and largest > 0.0:
        return v[:, -1] * np.sqrt(largest), True
    return None, False


def k_shape_of(spatial_shape):
    if len(spatial_shape) == 1:
        return (spatial_shape[0] // 2 + 1,)
    if len(spatial_shape) == 2:
        return (spatial_shape[0], spatial_shape[1] // 2 + 1)
    if len(spatial_shape) == 3:
        return (spatial_shape[0], spatial_shape[1], spatial_shape[2] // 2 + 1)
    raise NotImplementedError(f"make_elastic_updater not implemented for spatial_dim={len(spatial_shape)}")


class _DeviceElastic:
    def __init__(self, prec, spatial, bhat, l_vec, X0):
        _lib.require_gpu()
        h = C.c_void_p()
        b = np.ascontiguousarray(bhat, dtype=np.float64)
        _lib.check(_lib.lib().pf5d_elastic_create(prec, len(spatial), _lib.larr(spatial), _lib.ptr_f64(b),
                                                 _lib.darr(l_vec), _lib.darr(X0), C.byref(h)))
        self.handle, self.prec = h, prec

    def __call__(self, chemical_potentials, composition_field):
        M = Staged(chemical_potentials)
        X = Staged(composition_field)
        Xd = X.dev.contiguous()
        _lib.check(_lib.lib().pf5d_elastic_apply(self.handle, M.ptr(), row_stride(M.dev, 4),
                                                C.c_void_p(Xd.data_ptr()), _lib.stream_handle()))
        M.finish()

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib.lib().pf5d_elastic_destroy(h)


def make_elastic_updater(cfg, linear_elastic_coupling_matrix, elastic_kernel_r, field_shape, field_dtype,
                         fft_workers=1):
    """Return elastic_update(mu, X) adding the elastic potential in place (elastic.py:147-286).

    The rank-1 path (always taken for Vegard coupling) runs on the GPU.  The
    general path (non-rank-1 Lambda) is not reached by the reference driver and
    raises NotImplementedError here.
    """
    if not cfg.include_cubic_anisotropy or elastic_kernel_r is None:
        def elastic_update(mu, X):
            return None
        return elastic_update
    spatial = tuple(field_shape[:-1])
    k_shape = k_shape_of(spatial)
    assert elastic_kernel_r.shape == k_shape, (
        f"elastic_kernel_r shape {elastic_kernel_r.shape} does not match expected {k_shape}")
    l_vec, rank1 = rank_one_vector(linear_elastic_coupling_matrix)
    if not rank1:
        raise NotImplementedError("only the rank-1 (Vegard) elastic coupling runs on the device")
    prec = _lib.F32 if np.dtype(field_dtype) == np.float32 else _lib.F64
    X0 = np.asarray(cfg.initial_composition, dtype=np.float64)
    return _DeviceElastic(prec, spatial, elastic_kernel_r, l_vec.astype(field_dtype).astype(np.float64), X0)


def calculate_elastic_potential(reference_composition, composition_field, linear_elastic_coupling_matrix,
                                elastic_kernel_r, work_real=None, work_k=None, work_k_coupled=None, out=None,
                                fft_workers=1):
    """The general (non-rank-1) elastic path of elastic.py:114-140.  The Vegard coupling is
    always rank 1, so the reference driver never reaches it; it is not built here."""
    raise NotImplementedError("the general (non-rank-1) elastic potential is not implemented; "
                              "make_elastic_updater's rank-1 path covers Vegard coupling")
