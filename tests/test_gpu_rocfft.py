"""The hand-written Stockham FFT pair checked against rocFFT (through torch.fft on the
device) and against scipy.fft (the reference's own FFT, elastic.py:257-261), at every
length the workloads use: 2D sides 128/192/256/384/512 (in-plane grid, normal 0,0,1),
3D 64^3 / 96^3 / 128^3 (128^3 through the plane passes) and 256^3, and 1D serving
lengths.  psi = irfftn(Bhat rfftn(s)) of the engine's first step, s the misfit of the
initial field.  Bars: FP64 1e-12 normwise relative to both; FP32 1e-5 normwise
relative to the FP64 result and to rocFFT's own FP32 result.
"""
import numpy as np
import pytest

from helpers import rel_inf
from oracle import pf5d_oracle as O

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-5


def _problem(dim, side, X0=(0.1, 0.2, 0.2, 0.3), seed=0):
    from phasefield5d_amd.problem import build_problem
    from phasefield5d_amd.workloads import _mix_config, table_for
    return build_problem(_mix_config(dim, side, np.array(X0), 1e-15, 873, seed), table=table_for(873))


def _rocfft_psi(s, Bhat, dtype):
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    sd = torch.from_numpy(np.ascontiguousarray(s, dtype=dtype)).cuda()
    Bd = torch.from_numpy(np.ascontiguousarray(Bhat, dtype=dtype)).cuda()
    k = torch.fft.rfftn(sd, dim=tuple(range(s.ndim)))
    k *= Bd.to(tdt)
    return torch.fft.irfftn(k, s=s.shape, dim=tuple(range(s.ndim))).cpu().numpy().astype(np.float64)


def _engine_psi(pb, precision):
    from phasefield5d_amd.engine import Engine
    e = Engine(precision)
    slot = e.add_problem(pb)
    e.step(1)
    e.sync()
    info = e.launch_info()
    psi = e.field(slot, "psi").reshape(pb.shape)
    e.close()
    return psi, info


SHAPES = [(2, 128), (2, 192), (2, 256), (2, 384), (2, 512), (3, 64), (3, 96), (3, 128), (3, 256)]


@pytest.mark.parametrize("dim,side", SHAPES)
def test_engine_fft_vs_rocfft_and_scipy(dim, side):
    print(f"[fft] {dim}D side {side}", flush=True)
    pb = _problem(dim, side)
    s = O.misfit_scalar(pb.X_init, pb.X0, pb.l_vec)
    ref = O.elastic_psi(s, pb.Bhat)                         # scipy.fft, FP64
    roc64 = _rocfft_psi(s, pb.Bhat, np.float64)             # rocFFT, FP64
    psi64, info = _engine_psi(pb, "fp64")
    e_sc, e_roc = rel_inf(psi64, ref), rel_inf(psi64, roc64)
    print(f"[fft] fp64 vs scipy {e_sc:.2e} vs rocFFT {e_roc:.2e} (rocFFT vs scipy {rel_inf(roc64, ref):.2e}); "
          f"plane parts {info.get('fft_parts')}", flush=True)
    assert e_sc < TOL64 and e_roc < TOL64
    psi32, _ = _engine_psi(pb, "fp32")
    roc32 = _rocfft_psi(s - s.mean(), pb.Bhat, np.float32)
    e32, r32 = rel_inf(psi32, ref), rel_inf(roc32, ref)
    print(f"[fft] fp32 vs fp64 {e32:.2e} (rocFFT fp32 vs fp64 {r32:.2e}); fp32 vs rocFFT fp32 "
          f"{rel_inf(psi32, roc32):.2e}", flush=True)
    assert e32 < TOL32 and rel_inf(psi32, roc32) < TOL32


@pytest.mark.parametrize("n", [160, 200, 808, 1001, 2316, 3200])
def test_line_fft_vs_rocfft_1d(n):
    """1D serving lengths (mw 20-200 x ppw 8-16), including large prime factors, through
    the drop-in elastic updater."""
    from phasefield5d_amd.solver.elastic import build_khachaturyan_kernel, make_elastic_updater
    from phasefield5d_amd.solver.system import build_fourier_grid_1d_along_direction
    kg, kn = build_fourier_grid_1d_along_direction(n, 2e-9, [1, 0, 0])
    B = build_khachaturyan_kernel(kg, kn, 240.0, 150.0, 120.0)
    lih = np.array([0.086, 0.004, 0.008, 0.029])

    class Cfg:
        include_cubic_anisotropy = True
        initial_composition = np.array([0.1, 0.2, 0.2, 0.3])

    rng = np.random.default_rng(n)
    X = Cfg.initial_composition + 1e-3 * rng.normal(size=(n, 4))
    Lam = np.outer(lih, lih) * 7.1e-6
    from phasefield5d_amd.solver.elastic import rank_one_vector
    l, _ = rank_one_vector(Lam)
    mu = np.zeros((n, 4))
    make_elastic_updater(Cfg(), Lam, B, X.shape, np.float64)(mu, X)
    psi = mu[:, 0] / l[0]
    s = O.misfit_scalar(X, Cfg.initial_composition, l)
    ref = O.elastic_psi(s, B)
    roc = _rocfft_psi(s, B, np.float64)
    print(f"[fft] 1D {n}: vs scipy {rel_inf(psi, ref):.2e} vs rocFFT {rel_inf(psi, roc):.2e}", flush=True)
    assert rel_inf(psi, ref) < TOL64 and rel_inf(psi, roc) < TOL64
