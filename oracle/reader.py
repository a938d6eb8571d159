"""The reader's preprocessing of a spectrum before it is chunked, restated.

* ``mask_outliers``: QSOFile::maskOutliers (io/qso_file.cpp:189-230): the
  median and the scaled median absolute deviation of the unmasked pixels
  (1.4826 (1 + 1/sqrt(n)) MAD, at least 1), and every pixel outside
  [-1 - s, 1/F(lambda) + s] with s = factor * max(sigma, MAD) masked
  (F^-1 = exp(1.663e-3 (lambda / lambda_Lya)^4.107)).
* ``cut_boundary``: QSOFile::cutBoundary (io/qso_file.cpp:159-187) to the
  redshift range of the bins, then the masked-boundary cut
  (io/qso_file.cpp:139-157).
* ``read_spectrum``: OneQSOEstimate::_readQsoFile for binary files with no
  flux conversion and no resolution matrix (core/one_qso_estimate.cpp:52-95):
  read, mask outliers, cut; None where the reference throws (low SNR, empty,
  fewer than 100 unmasked pixels).

Test infrastructure only (see oracle/__init__.py).  Pins
(tests/test_oracle_golden.py): the reference's test spectrum passes whole
and its chunk reproduces the golden Fisher matrix; constructed outliers at
either side of each threshold (MAD floor of 1, per-pixel sigma, the mean
flux bound); the redshift and masked-boundary cuts against their closed
forms.
"""
from __future__ import annotations

import math

import numpy as np

from .bins import Bins, LYA_REST

MIN_PIXELS_IN_FILE = 100


def _median(a: np.ndarray) -> float:
    s = np.sort(np.asarray(a, float))
    n = s.size
    return float((s[n // 2 - 1] + s[n // 2]) / 2) if n % 2 == 0 else float(s[n // 2])


def mask_outliers(wave, delta, ivar, factor: float = 5.0):
    """Returns (delta, ivar, number masked), copies."""
    wave = np.asarray(wave, float)
    delta, ivar = np.array(delta, float), np.array(ivar, float)
    good = ivar != 0
    vals = delta[good]
    med = _median(vals)
    mad = 1.4826 * (1.0 + 1.0 / math.sqrt(int(good.sum()))) * _median(np.abs(vals - med))
    mad = max(1.0, mad)
    n = 0
    for i in np.flatnonzero(good):
        s = factor * max(1.0 / math.sqrt(ivar[i]), mad)
        fup = math.exp(1.663e-03 * math.pow(wave[i] / LYA_REST, 4.107))
        if delta[i] < -1.0 - s or delta[i] > fup + s:
            delta[i] = 0.0
            ivar[i] = 0.0
            n += 1
    return delta, ivar, n


def cut_boundary(bins: Bins, wave, delta, ivar):
    """The slice [a, b) the cut keeps, or None if it leaves nothing."""
    wave, ivar = np.asarray(wave, float), np.asarray(ivar, float)
    l1, l2 = LYA_REST * (1 + bins.z_lower_edge), LYA_REST * (1 + bins.z_upper_edge)
    a, b = int(np.searchsorted(wave, l1, "left")), int(np.searchsorted(wave, l2, "right"))
    if a == wave.size or b == 0:
        return None
    nz = np.flatnonzero(ivar[a:b] != 0)
    if nz.size == 0:
        return None
    return a + int(nz[0]), a + int(nz[-1]) + 1


def read_spectrum(bins: Bins, wave, delta, ivar, snr: float = 1.0, min_snr: float = 0.0):
    """(wave, delta, ivar) as the chunks see them, or None."""
    if snr < min_snr:
        return None
    delta, ivar, _ = mask_outliers(wave, delta, ivar)
    cut = cut_boundary(bins, wave, delta, ivar)
    if cut is None:
        return None
    a, b = cut
    w, d, iv = np.asarray(wave, float)[a:b], delta[a:b], ivar[a:b]
    if np.count_nonzero(iv) < MIN_PIXELS_IN_FILE:
        return None
    return w, d, iv
