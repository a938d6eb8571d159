"""What OneDQuadraticPowerEstimate::iterate does with the accumulated sums
after its chunk loop (core/quadratic_estimate.cpp:448-521): the 0.5 scaling
and symmetrisation of the Fisher sum (:456-457), mxhelp::stableInvertSym
(mathtools/matrix_helper.cpp:219-230,294-358), the power estimate
theta = 0.5 F_solver (d - b - t) (:290-318), sigma = sqrt(diag F^-1)
(:717-740), and the convergence test (:548-614).  T x T host arithmetic in
FP64 (T = 455 for the DESI-like binning), once per iteration.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import lu_factor, lu_solve


def normalize_fisher(fisher_sum: np.ndarray) -> np.ndarray:
    f = 0.5 * fisher_sum
    iu = np.triu_indices(f.shape[0], 1)
    f.T[iu] = f[iu]
    return f


def _inv(a):
    lu = lu_factor(a, check_finite=True)
    return lu_solve(lu, np.eye(a.shape[0]))


def stable_invert_sym(f: np.ndarray):
    """Returns (solver, inverse, damp, dof)."""
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
        sinv = _inv(s)
        a = s @ s
        a[np.diag_indices(n)] += damp
        solver = _inv(a) @ s
    else:
        solver = _inv(s)
        sinv = solver.copy()
    for i in empty:
        solver[i, i] = 0.0
        sinv[i, i] = 0.0
    return solver, sinv, damp, n - len(empty)


def power_estimates(fisher_sum: np.ndarray, dbt_sum: np.ndarray) -> dict:
    fs = normalize_fisher(fisher_sum)
    solver, sinv, damp, dof = stable_invert_sym(fs)
    weighted = np.array([0.5 * solver @ dbt_sum[i] for i in range(3)])
    theta = weighted[0] - weighted[1] - weighted[2]
    with np.errstate(invalid="ignore"):
        sigma = np.sqrt(np.diag(sinv))
    return dict(fisher=fs, solver=solver, inverse=sinv, theta=theta, sigma=sigma,
                weighted=weighted, damp=damp, dof=dof)
