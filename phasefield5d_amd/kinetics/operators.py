"""Linear kinetic operator B(k) = M (H + k^2 kappa) and its growth-rate scan.

Restates phasefield5d/kinetics/operators.py:8-87.  The reference diagonalises
B(k) one wavenumber at a time in a Python loop; here the 10,000 4x4 problems
are handed to LAPACK as one stacked ``numpy.linalg.eig`` call (the same geev
routine per matrix), and the argsort/argmax selection is vectorised.  The
selection rule is unchanged: per k, sort eigenvalues by real part, growth
rate r = -k^2 lambda, take the largest r over all k, first occurrence wins.
"""
import numpy as np
import scipy.linalg as la


def calculate_wavenumber_cutoff(hessian, kappa_matrix):
    """Largest unstable wavenumber k* = sqrt(-lambda_min(H, kappa)); 0 if stable."""
    if np.all(kappa_matrix == 0):
        lam = np.linalg.eigvalsh(hessian)
    else:
        lam = la.eigvalsh(hessian, kappa_matrix)
    return float(np.sqrt(-lam[0])) if lam[0] < 0 else 0.0


def calculate_kinetic_operators(hessian, mobility_matrix, kappa_matrix, n_k=50):
    """Sample B(k) on k in [0.001 k*, 0.999 k*]; NaN arrays when stable."""
    n = hessian.shape[0]
    k_cut = calculate_wavenumber_cutoff(hessian, kappa_matrix)
    if k_cut == 0:
        return np.full(n_k, np.nan), np.full((n_k, n, n), np.nan)
    ks = np.linspace(0.001 * k_cut, 0.999 * k_cut, n_k)
    ops = np.stack([mobility_matrix @ (hessian + k**2 * kappa_matrix) for k in ks])
    return ks, ops


def calculate_kinetic_eigenvalues_vectors(wavenumbers, kinetic_operators, max_tag=False):
    """Growth-rate scan.  ``max_tag=True`` -> (rate, k, eigenvalue, mode)."""
    ops = np.asarray(kinetic_operators)
    n = ops.shape[-1] if not np.isnan(ops).all() else 4
    if np.isnan(ops).all():
        if max_tag:
            return np.nan, np.nan, np.nan, np.full(n, np.nan)
        return {"growth_rates": [], "wavenumbers": [], "eigenvalues": [], "eigenvectors": []}
    ks = np.asarray(wavenumbers)
    finite = np.isfinite(ops).reshape(len(ops), -1).all(axis=1)
    ks, ops = ks[finite], ops[finite]
    vals, vecs = np.linalg.eig(ops)
    order = np.argsort(np.real(vals), axis=-1)
    vals = np.take_along_axis(vals, order, axis=-1)
    vecs = np.take_along_axis(vecs, order[:, None, :], axis=-1)
    rates = -(ks[:, None] ** 2) * vals
    if not max_tag:
        return {"growth_rates": list(rates), "wavenumbers": list(ks),
                "eigenvalues": list(vals), "eigenvectors": list(vecs)}
    if len(ks) == 0:
        return np.nan, np.nan, np.nan, np.full(n, np.nan)
    # per-k argmax over the sorted spectrum (numpy argmax on complex compares
    # lexicographically (real, imag), exactly as the per-k loop does)
    per_k = np.argmax(rates, axis=-1)
    best = rates[np.arange(len(ks)), per_k]
    # strict '>' update across k in the reference == first occurrence of the max real part
    kbest = int(np.argmax(np.real(best)))
    i = per_k[kbest]
    return float(np.real(best[kbest])), ks[kbest], vals[kbest, i], vecs[kbest, :, i]
