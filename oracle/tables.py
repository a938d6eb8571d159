"""The S/Q lookup-table generator, restated (SQLookupTable::computeTables,
core/sq_table.cpp:201-323).

* Derivative table: Q_kn(v) = (1/pi) * int_{k1}^{k2} W(k)^2 cos(k v) dk with
  W(k) = sinc(k dv/2) exp(-(k R)^2/2), R the 1-sigma resolution in km/s
  (core/fiducial_cosmology.cpp:305-331), by QUADPACK's QAWO with
  epsabs=1e-13, epsrel=1e-7 (mathtools/fourier_integrator.cpp:85-98).  GSL's
  gsl_integration_qawo is a port of the same QUADPACK routine; scipy's
  ``quad(weight='cos')`` calls the original (dqawoe), so this is the same
  published algorithm.
* Signal table: field_k = W^2 P_fid(k, z) on a real-FFT grid of
  n = 2^(ceil(log2(4 Nv)) + 1) points with spacing dx = (Lv/(Nv-1))/2,
  transformed c2r and divided by the grid length; every 2nd point is kept
  (core/sq_table.cpp:221-227,296-311; mathtools/real_field.cpp:40-92).

Pins: tests/truth/signal_R70000_dv20.0.fits (both tables) at the reference
test's tolerance, rel 1e-5 / abs 1e-8 (tests/compareTestResults.py:8-9,
80-100), in tests/test_oracle_golden.py::test_tables_against_golden_fits.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import integrate

from .bins import MY_PI, ONE_SIGMA_2_FWHM, SPEED_OF_LIGHT, Bins, fiducial_pd13


def window_sq(k, dv_kms: float, R_kms: float):
    k = np.asarray(k, dtype=np.float64)
    x = k * dv_kms / 2.0
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.where(x == 0, 1.0, np.sin(x) / np.where(x == 0, 1.0, x))
    kr = k * R_kms
    w = s * np.exp(-kr * kr / 2.0)
    return w * w


def derivative_table_qawo(bins: Bins, R_fwhm: int, dv_kms: float, nv: int, v_length: float):
    rkms = SPEED_OF_LIGHT / R_fwhm / ONE_SIGMA_2_FWHM
    dx = 0.0 if nv == 1 else v_length / (nv - 1.0)
    v = dx * np.arange(nv)
    out = np.zeros((bins.nk, nv))
    f = lambda k: float(window_sq(k, dv_kms, rkms)) / MY_PI
    for kn in range(bins.nk):
        k1, k2 = bins.kband_edges[kn], bins.kband_edges[kn + 1]
        for j in range(nv):
            out[kn, j] = integrate.quad(f, k1, k2, weight="cos", wvar=v[j],
                                        epsabs=1e-13, epsrel=1e-7, limit=3000)[0]
    return out


def signal_table_fft(bins: Bins, R_fwhm: int, dv_kms: float, nv: int, v_length: float,
                     nz: int, pd13: tuple):
    rkms = SPEED_OF_LIGHT / R_fwhm / ONE_SIGMA_2_FWHM
    itp_dv = 0.0 if nv == 1 else v_length / (nv - 1.0)
    nrft = int(2.0 ** (math.ceil(math.log2(2 * 2 * nv)) + 1))
    dxf = itp_dv / 2
    length = dxf * nrft
    kf = 2 * MY_PI / length
    k = np.arange(nrft // 2 + 1) * kf
    z1 = bins.zbin_centers[0] - bins.z_bin_width
    zlen = bins.z_bin_width * (bins.nz + 1)
    dz = 0.0 if nz == 1 else zlen / (nz - 1.0)
    w2 = window_sq(k, dv_kms, rkms)
    out = np.zeros((nv, nz))
    for iz in range(nz):
        z = z1 + dz * iz
        field_k = w2 * fiducial_pd13(k, z, *pd13)
        field_x = np.fft.irfft(field_k, nrft) * nrft / length
        out[:, iz] = field_x[2 * np.arange(nv)]
    return out
