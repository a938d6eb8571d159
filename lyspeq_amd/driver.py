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


def _smoothing_spline(x, y, z, w, s):
    """scipy's SmoothBivariateSpline(x, y, z, w=w, s=s) (bicubic, FITPACK
    surfit), except for the workspace: scipy retries surfit once with the
    lwrk2 that the first call asks for (ier > 10), and when the retry asks
    for more it warns and keeps the unfinished fit.  Here the workspace grows
    until surfit completes, so the fit is the one the routine defines."""
    import warnings
    from scipy.interpolate import BivariateSpline
    from scipy.interpolate import _dfitpack as dfitpack
    x, y, z, w = (np.ascontiguousarray(a, dtype=float) for a in (x, y, z, w))
    lwrk2 = 1
    for _ in range(16):
        nx, tx, ny, ty, c, fp, _, ier = dfitpack.surfit_smth(x, y, z, w, kx=3, ky=3, s=float(s), eps=1e-16,
                                                            lwrk2=lwrk2)
        if ier <= 10:
            break
        lwrk2 = ier
    if ier not in (0, -1, -2):
        warnings.warn(f"surfit: ier={ier} (fp={fp:g}, s={s:g})", stacklevel=3)
    spl = BivariateSpline._from_tck((tx[:nx], ty[:ny], c[:(nx - 4) * (ny - 4)], 3, 3))
    spl.fp, spl.ier = fp, ier   # surfit's weighted residual sum and return code
    return spl


def smooth_power(theta: np.ndarray, sigma: np.ndarray, kcenters: np.ndarray, zcenters: np.ndarray,
                 pfid: np.ndarray, interp_log: bool = False) -> np.ndarray:
    """The fits the next iteration uses, as OneDQuadraticPowerEstimate::
    _smoothPowerSpectra makes them (core/quadratic_estimate.cpp:353-396):
    writeSpectrumEstimates (:624-650) lists z ("%.3lf"), k, P = theta + P_fid
    and e = sqrt((F^-1)_ii) ("%e") per band, z-major; py/smbivspline.py fits
    scipy's SmoothBivariateSpline(z, k, P, w=1/e, s=len(e)), or with
    interp_log (SmoothLogkLogP) one in (ln(1+z), ln k, ln P) with weights
    P/e over the bands with P, e > 0, and evaluates it at every band;
    _readScriptOutput (:340-348) subtracts P_fid again.  Bands with e <= 0
    (no data), whose weight the script would make infinite, are left out of
    the fit here.  Where scipy's one retry leaves surfit short of workspace
    (ier > 10: an unfinished fit and a warning in the reference's run), the
    workspace is grown until the fit completes (_smoothing_spline).
    kcenters (nk) and zcenters (nz) give the band grid; theta, sigma and pfid
    have nz * nk values."""
    nk, nz = len(kcenters), len(zcenters)
    assert theta.size == sigma.size == pfid.size == nk * nz
    zz = np.repeat(np.asarray(zcenters, float), nk)
    kk = np.tile(np.asarray(kcenters, float), nz)
    # the values the script reads back from the text file
    z = np.array([float(f"{v:.3f}") for v in zz])
    k = np.array([float(f"{v:e}") for v in kk])
    p = np.array([float(f"{v:e}") for v in theta + pfid])
    e = np.array([float(f"{v:e}") for v in np.nan_to_num(sigma, nan=0.0)])
    return smooth_values(z, k, p, e, interp_log) - pfid


def smooth_values(z, k, p, e, interp_log: bool = False) -> np.ndarray:
    """py/smbivspline.py's fit and evaluation on the values of its input
    file (z, k, P, e per band): the weighted smoothing spline with s = the
    number of points fitted, in (z, k, P) with weights 1/e, or with
    interp_log in (ln(1+z), ln k, ln P) with weights P/e over the bands with
    P, e > 0; bands with e <= 0 are left out of the fit."""
    z, k, p, e = (np.asarray(a, float) for a in (z, k, p, e))
    if interp_log:
        m = (p > 0) & (e > 0)
        spl = _smoothing_spline(np.log(1 + z[m]), np.log(k[m]), np.log(p[m]), p[m] / e[m], int(m.sum()))
        return np.exp(spl(np.log(1 + z), np.log(k), grid=False))
    m = e > 0
    spl = _smoothing_spline(z[m], k[m], p[m], 1.0 / e[m], int(m.sum()))
    return spl(z, k, grid=False)
