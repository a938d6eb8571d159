"""The Newton-Raphson driver's use of F, d, b, t, restated.

* ``normalize_fisher``: cblas_dscal(0.5) + copyUpperToLower
  (core/quadratic_estimate.cpp:456-457, mathtools/matrix_helper.cpp:81-94).
* ``stable_invert_sym``: mxhelp::stableInvertSym
  (mathtools/matrix_helper.cpp:219-230,294-358): empty diagonals set to 1,
  the SVD jump test (a singular value whose predecessor is > 8x larger sets
  the damping), LU inverse, optional damped solver (S S + damp I)^-1 S, and
  the empty diagonals zeroed again.
* ``power_estimates``: theta = 0.5 * solver * (d - b - t) and the Fisher
  weighted d/b/t (core/quadratic_estimate.cpp:290-318); sigma =
  sqrt(diag(inverse)) (core/quadratic_estimate.cpp:717-740).
* ``has_converged``: the relative-change statistics and the two chi^2/dof
  tests (core/quadratic_estimate.cpp:548-614).

Pins: tests/truth/test_it1_quadratic_power_estimate_detailed.dat (ThetaP,
ErrorP, Fisher-weighted d/b/t) and test_it1_inversefisher_matrix.dat.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import lapack


def normalize_fisher(fisher_sum: np.ndarray) -> np.ndarray:
    f = 0.5 * fisher_sum
    iu = np.triu_indices(f.shape[0], 1)
    f.T[iu] = f[iu]
    return f


def _lu_inverse(a: np.ndarray) -> np.ndarray:
    lu, piv, info = lapack.dgetrf(a)
    if info != 0:
        raise np.linalg.LinAlgError("ERROR in LU decomposition.")
    inv, info = lapack.dgetri(lu, piv)
    if info != 0:
        raise np.linalg.LinAlgError("ERROR in LU invert.")
    return inv


def stable_invert_sym(f: np.ndarray):
    """Returns (solver, inverse, damped, damp, dof)."""
    s = f.copy()
    n = s.shape[0]
    empty = [i for i in range(n) if s[i, i] == 0]
    for i in empty:
        s[i, i] = 1.0
    sv = np.linalg.svd(s, compute_uv=False)
    damp = 0.0
    for i in range(n - 1, 0, -1):
        if sv[i - 1] / sv[i] > 8.0:
            damp = sv[i]
        else:
            break
    if damp != 0:
        sinv = _lu_inverse(s)
        a = s @ s
        a[np.diag_indices(n)] += damp
        solver = _lu_inverse(a) @ s
    else:
        solver = _lu_inverse(s)
        sinv = solver.copy()
    for i in empty:
        solver[i, i] = 0.0
        sinv[i, i] = 0.0
    return solver, sinv, damp != 0, damp, n - len(empty)


def power_estimates(fisher_sum: np.ndarray, dbt_sum: np.ndarray):
    """From the accumulated (unhalved, upper) Fisher sum and d/b/t sums to
    theta, sigma and the Fisher-weighted d/b/t."""
    fs = normalize_fisher(fisher_sum)
    solver, sinv, damped, damp, dof = stable_invert_sym(fs)
    weighted = np.array([0.5 * solver @ dbt_sum[i] for i in range(3)])
    theta = weighted[0] - weighted[1] - weighted[2]
    sigma = np.sqrt(np.diag(sinv))
    return dict(fisher=fs, solver=solver, inverse=sinv, theta=theta, sigma=sigma,
                weighted=weighted, damped=damped, damp=damp, dof=dof)


def has_converged(current, previous, inverse, fisher, dof, eps):
    """Returns (abs_mean, abs_max, chi2_diag, chi2_full, converged)."""
    current, previous = np.asarray(current, float), np.asarray(previous, float)
    p1, p2 = np.abs(current), np.abs(previous)
    keep = ~((p1 == 0) & (p2 == 0))
    r = np.abs(p1 - p2)[keep] / np.maximum(p1, p2)[keep]
    abs_mean = float(np.sum(r / dof))
    abs_max = float(r.max()) if r.size else 0.0
    t = previous - current
    e = np.diag(inverse)
    ok = e > 0
    chi2_diag = float(np.sqrt(np.sum(t[ok] ** 2 / e[ok]) / dof))
    up = np.triu(fisher)
    sym = up + np.triu(up, 1).T          # dsymv reads the upper triangle
    chi2_full = float(np.sqrt(abs(t @ sym @ t) / dof))
    return abs_mean, abs_max, chi2_diag, chi2_full, bool(chi2_diag < eps or chi2_full < eps)
