"""The reference's FastBootstrap, restated (test infrastructure only: the
product path never imports this module).

* ``collapse``: OneQSOEstimate::collapseBootstrap
  (core/one_qso_estimate.cpp:200-228): istart = min fisher_index_start over
  the spectrum's chunks, ndim = max(fisher_index_start + N_Q) - istart, and
  theta_q = sum over the chunks, in order, of (d - b) - t at their offset.
* ``realizations``: PoissonBootstrapper::_fastBootstrap with
  OneQSOEstimate::addBootPowerOnly (core/bootstrapper.hpp:183-221,
  core/one_qso_estimate.hpp:55-60): for each spectrum in queue order, nboots
  weights p (the next draws of one PoissonRNG) and R[:, istart:istart+ndim] +=
  p theta_q^T (cblas_dger).
* ``mean_covariance``: _calcuateMean and _calcuateCovariance
  (core/bootstrapper.hpp:282-375): mean = (sum of R[b]) * (1/nboots); cov =
  sum of (R[b] - mean)(R[b] - mean)^T, times 1/(nboots - 1).

* ``collapse_fisher`` and ``complete_realizations``: the complete form
  (FastBootstrap off), a Cholesky solve per realisation.
* ``median`` and ``median_covariance``: _medianBootstrap, _calcuateMedian and
  _calcuateMadCovariance (core/bootstrapper.hpp:301-346,376-390) with
  stats::medianOfUnsortedVector / medianOfSortedArray (mathtools/stats.cpp:
  8-28): a band's median over the realisations (middle value, or the half-sum
  of the two middle values for even counts), the deviations from it, and per
  band pair j >= i the median of the products of the deviations
  (vector_multiply, mathtools/matrix_helper.hpp:47-52), mirrored.

Pins (tests/test_bootstrap_oracle.py).  Fast form: with unit weights every
realisation is the accumulated d - b - t sum of the included chunks and the
covariance vanishes; with Poisson(1) weights the covariance tends to sum_q theta_q
theta_q^T (independent unit-variance weights), which a dropped mean
subtraction or a wrong normalisation fails; a brute-force element loop on a
tiny case.  Complete form: with unit weights every realisation is the
driver's power estimate (F_sum)^-1 (d - b - t), which the driver restatement
computes by its own LU path (pinned by the golden iteration files); a
realisation whose F_b does not factorise is dropped.  Median form: a brute
loop over sorted() lists on a tiny case; numpy's median (the same
definition); for odd counts the diagonal is the squared median absolute
deviation (scipy.stats.median_abs_deviation, scale 1), which a dropped median
subtraction fails.  The reference's Poisson
draws themselves come from the C++ standard library; no reference fixture
holds them.
"""
from __future__ import annotations

from typing import List, Sequence, Tuple

import numpy as np


def collapse(chunks: Sequence[Tuple[int, int, np.ndarray]]):
    """chunks: (fisher_index_start, n_q, dbt 3 x n_q) in chunk order.
    Returns (istart, theta of length ndim)."""
    istart = min(f for f, _, _ in chunks)
    iend = max(f + n for f, n, _ in chunks)
    theta = np.zeros(iend - istart)
    for f, n, dbt in chunks:
        v = (dbt[0] - dbt[1]) - dbt[2]
        theta[f - istart:f - istart + n] += v
    return istart, theta


def realizations(P: np.ndarray, collapsed: List[Tuple[int, np.ndarray]], T: int) -> np.ndarray:
    """P: nqso x nboots weights, collapsed[q] = (istart, theta_q)."""
    nboots = P.shape[1]
    R = np.zeros((nboots, T))
    for q, (istart, theta) in enumerate(collapsed):
        R[:, istart:istart + theta.size] += np.outer(P[q].astype(np.float64), theta)
    return R


def mean_covariance(R: np.ndarray):
    n = R.shape[0]
    mean = R.sum(axis=0) * (1.0 / n)
    v = R - mean
    cov = (v.T @ v) * (1.0 / (n - 1))
    return mean, cov


def collapse_fisher(chunks: Sequence[Tuple[int, int, np.ndarray]]):
    """The spectrum's Fisher block (collapseBootstrap without FastBootstrap,
    core/one_qso_estimate.cpp:224-234): chunks are (fisher_index_start, n_q,
    F upper n_q x n_q); returns (istart, ndim x ndim upper)."""
    istart = min(f for f, _, _ in chunks)
    iend = max(f + n for f, n, _ in chunks)
    F = np.zeros((iend - istart, iend - istart))
    for f, n, fc in chunks:
        F[f - istart:f - istart + n, f - istart:f - istart + n] += np.triu(fc)
    return istart, F


def complete_realizations(W: np.ndarray, collapsed, fishers, T: int, return_systems: bool = False):
    """PoissonBootstrapper::_oneBoot (core/bootstrapper.hpp:223-262) with
    OneQSOEstimate::addBoot (core/one_qso_estimate.hpp:43-53) and
    LAPACKE_safeSolveCho (mathtools/matrix_helper.cpp:385-396).  W is
    nboots x nqso (one weight per spectrum per realisation).  Returns the
    valid realisations in order and the validity flags, and with
    return_systems the (F_b, rhs) each valid one solves."""
    from scipy.linalg import cho_factor, cho_solve
    out, valid, systems = [], [], []
    for w in W:
        F = np.zeros((T, T))
        r = np.zeros(T)
        for q, p in enumerate(w):
            if p == 0:
                continue
            istart, th = collapsed[q]
            _, Fq = fishers[q]
            n = th.size
            F[istart:istart + n, istart:istart + n] += p * Fq
            r[istart:istart + n] += p * th
        S = np.triu(F) + np.triu(F, 1).T
        empty = np.diag(S) == 0
        S[empty, empty] = 1.0
        try:
            x = cho_solve(cho_factor(S, lower=False), r)
        except np.linalg.LinAlgError:
            valid.append(False)
            continue
        x[empty] = 0.0
        out.append(x)
        valid.append(True)
        systems.append((S, r))
    if return_systems:
        return np.array(out).reshape(-1, T), np.array(valid), systems
    return np.array(out).reshape(-1, T), np.array(valid)


def _median_rows(X: np.ndarray) -> np.ndarray:
    """Median of each row: the sorted middle value, or (s[n/2-1] + s[n/2]) / 2."""
    n = X.shape[-1]
    if n % 2:
        return np.partition(X, n // 2, axis=-1)[..., n // 2]
    p = np.partition(X, (n // 2 - 1, n // 2), axis=-1)
    return (p[..., n // 2 - 1] + p[..., n // 2]) / 2


def median(R: np.ndarray) -> np.ndarray:
    """R: nboots x T realisations; per band median."""
    return _median_rows(np.ascontiguousarray(R.T))


def median_covariance(R: np.ndarray):
    """(median T, median covariance T x T) of realisations R (nboots x T)."""
    X = np.ascontiguousarray(R.T, dtype=np.float64)
    med = _median_rows(X)
    X = X - med[:, None]
    T = X.shape[0]
    mad = np.zeros((T, T))
    for i in range(T):
        row = _median_rows(X[i][None, :] * X[i:])
        mad[i, i:] = row
        mad[i:, i] = row
    return med, mad
