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

Pins (tests/test_bootstrap_oracle.py): with unit weights every realisation is
the accumulated d - b - t sum of the included chunks and the covariance
vanishes; with Poisson(1) weights the covariance tends to sum_q theta_q
theta_q^T (independent unit-variance weights), which a dropped mean
subtraction or a wrong normalisation fails; a brute-force element loop on a
tiny case.  The reference's Poisson draws themselves come from the C++
standard library; no reference fixture holds them.
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
