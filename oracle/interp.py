"""Uniform-grid interpolation with the reference's index arithmetic.

``evaluate_1d`` follows DiscreteInterpolation1D::evaluate
(mathtools/discrete_interpolation.cpp:61-72): xx = (x - x1)/dx, n = (int)xx
(truncation toward zero), n clamped to [0, N-2], y[n](1-d) + y[n+1]d with
d = xx - n, which extrapolates linearly past either end.

``evaluate_2d`` follows DiscreteInterpolation2D::evaluate
(mathtools/discrete_interpolation.cpp:221-242) with x the fast axis of the
z array and the same clamp rule on both axes; the four products are summed in
the reference's order.

Pins: the golden signal_matrix.txt and q0_matrix.txt (all 488^2 entries) and
closed forms (exact reproduction at grid nodes, linearity between nodes,
extrapolation beyond the ends) in tests/test_oracle_golden.py::test_interp_closed_forms.
"""
from __future__ import annotations

import numpy as np


def _index(xx: np.ndarray, n: int) -> np.ndarray:
    idx = np.trunc(xx).astype(np.int64)       # C (int) conversion
    idx = np.where(idx < 0, 0, idx)
    idx = np.where(idx >= n - 1, n - 2, idx)
    return idx


def evaluate_1d(x, x1: float, dx: float, y: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.float64)
    xx = (x - x1) / dx
    n = _index(xx, y.size)
    dn = xx - n
    return y[n] * (1 - dn) + y[n + 1] * dn


def evaluate_2d(x, y, x1: float, dx: float, y1: float, dy: float,
                z: np.ndarray, nx: int, ny: int) -> np.ndarray:
    """z is the flat table with z[ix + nx*iy]."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    xx = (x - x1) / dx
    yy = (y - y1) / dy
    ix = _index(xx, nx)
    iy = _index(yy, ny)
    ind = ix + nx * iy
    dnx = xx - ix
    dny = yy - iy
    return (z[ind] * (1 - dnx) * (1 - dny)
            + z[ind + 1] * dnx * (1 - dny)
            + z[ind + nx] * (1 - dnx) * dny
            + z[ind + nx + 1] * dnx * dny)
