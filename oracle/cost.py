"""The reference's relative compute-time model for load balancing, restated.

* ``chunk_cost``: Chunk::getComputeTimeEst (core/chunk_estimate.cpp:297-382):
  the masked-boundary cut of the pixel range, the short-chunk test, the
  redshift bins and the triangular extension, KBIN_UPP from the median pixel
  width, and the NERSC Perlmutter ("-c 2") per-call costs dgemm = 19 x^3,
  dsymv = 0.45 x^2, ddot = 1.0 x^2 with x = N / 100, combined as
  n (dgemm + ddot (n + 2)) + (n + 1) dsymv [+ dgemm + n ddot with S_fid]
  [+ 11 ddot n (os (os + 1), or 2) with resolution matrices], n the stored Q
  count (FISHER_OPTIMIZATION off, the configure default).
* ``spectrum_cost``: OneQSOEstimate::getComputeTimeEst
  (core/one_qso_estimate.cpp:128-166): the redshift boundary cut, the
  100-pixel minimum, decideIndices and the sum over the chunks.

Test infrastructure only (see oracle/__init__.py).  Pins: a hand-evaluated
chunk and the scaling of each term with N (tests/test_dropin_cpp.py).
"""
from __future__ import annotations

import math

import numpy as np

from .bins import Bins, LYA_REST, MY_PI, SPEED_OF_LIGHT, TRIANGULAR
from .chunk import decide_indices, median_sorted

MIN_PIXELS_IN_FILE = 100


def chunk_cost(bins: Bins, wave, ivar, i1: int, i2: int, use_rmat: bool = False, oversampling: int = -1) -> float:
    w = np.asarray(wave, float)[i1:i2]
    iv = np.asarray(ivar, float)[i1:i2]
    nz_idx = np.flatnonzero(iv != 0)
    if nz_idx.size == 0:          # the QSOFile constructor throws; the estimate is 0
        return 0.0
    w, iv = w[nz_idx[0]:nz_idx[-1] + 1], iv[nz_idx[0]:nz_idx[-1] + 1]
    N = w.size
    real = int(np.count_nonzero(iv))
    lenv = SPEED_OF_LIGHT * math.log(w[-1] / w[0])
    if real < int(lenv / bins.max_pixel_length_v) or lenv < bins.min_forest_length_v:
        return 0.0
    z1, z2 = w[0] / LYA_REST - 1, w[-1] / LYA_REST - 1
    zlo, zup = bins.find_redshift_bin(z1), bins.find_redshift_bin(z2)
    if zlo > bins.nz - 1 or zup < 0:
        return 0.0
    nq = zup - zlo + 1
    fidx = bins.nk * zlo
    if bins.method == TRIANGULAR:
        if zlo >= 0 and z1 < bins.zbin_centers[zlo] and zlo != 0:
            nq += 1
            fidx -= bins.nk
        if zup < bins.nz and bins.zbin_centers[zup] < z2 and zup < bins.nz - 1:
            nq += 1
    nq *= bins.nk
    dv = median_sorted(np.sort(np.log(w[1:] / w[:-1]) * SPEED_OF_LIGHT))
    kncut = bins.max_k_index(MY_PI / dv)
    n = sum(1 for i in range(nq) if (i + fidx) % bins.nk < kncut)
    x = N / 100.0
    dgemm, dsymv, ddot = 19.0 * x ** 3, 0.45 * x ** 2, 1.0 * x ** 2
    res = n * (dgemm + ddot * (n + 2)) + (n + 1) * dsymv
    if not bins.turn_off_sfid:
        res += dgemm + n * ddot
    if use_rmat:
        extra = 11 * ddot * n
        extra *= oversampling * (oversampling + 1) if oversampling > 0 else 2
        res += extra
    return res


def spectrum_cost(bins: Bins, wave, ivar, use_rmat: bool = False, oversampling: int = -1):
    """(cost, zbin) of one spectrum after the reader's redshift boundary cut."""
    wave, ivar = np.asarray(wave, float), np.asarray(ivar, float)
    l1, l2 = LYA_REST * (1 + bins.z_lower_edge), LYA_REST * (1 + bins.z_upper_edge)
    a, b = int(np.searchsorted(wave, l1, "left")), int(np.searchsorted(wave, l2, "right"))
    if a == wave.size or b == 0:
        return 0.0, -1
    w, iv = wave[a:b], ivar[a:b]
    nz_idx = np.flatnonzero(iv != 0)
    if nz_idx.size == 0:
        return 0.0, -1
    w, iv = w[nz_idx[0]:nz_idx[-1] + 1], iv[nz_idx[0]:nz_idx[-1] + 1]
    if np.count_nonzero(iv) < MIN_PIXELS_IN_FILE:
        return 0.0, -1
    zbin = bins.find_redshift_bin(median_sorted(w) / LYA_REST - 1)
    idx = decide_indices(w, bins)
    return sum(chunk_cost(bins, w, iv, idx[k], idx[k + 1], use_rmat, oversampling)
               for k in range(len(idx) - 1)), zbin
