"""The reference's end-to-end test case (tests/input/test.config), as data.

One 488-pixel chunk (R=70000, dv=20 km/s, z 1.961-2.059), 3 linear + 2 log k
bins, 2 triangular z bins centred at 1.8 and 2.0, continuum marginalisation
off, fiducial signal on, one iteration with theta = 0.
"""
import os

import numpy as np

from lyspeq_amd import sqio

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

BINS_KW = dict(k0=0.0, nklin=3, dklin=7e-4, nklog=2, dklog=0.5, z0=1.8,
               z_bin_width=0.2, nz=2, klast=-1.0, method=1,
               cont_logl_order=-1, cont_lam_order=-1, turn_off_sfid=False,
               max_forest_length_v=14000.0, number_of_chunks=1,
               max_pixel_length_v=150.0)
PD13 = (0.06, -2.6, -0.1, 1.0, 0.0, 0.0)


def table():
    return sqio.read_sq_table(os.path.join(GOLDEN, "signal_R70000_dv20.0.fits"))


def spectrum():
    return sqio.read_bq_dat(os.path.join(GOLDEN, "desilite_seed332298_id0_0_z2.4_lognormal.dat"))


def detailed_rows():
    with open(os.path.join(GOLDEN, "test_it1_quadratic_power_estimate_detailed.dat")) as f:
        lines = [l for l in f if not l.startswith("#")]
    return np.loadtxt(lines[1:])


def golden_fisher():
    return sqio.read_text_matrix(os.path.join(GOLDEN, "test_it1_fisher_matrix.dat"))


def golden_inverse_fisher():
    return sqio.read_text_matrix(os.path.join(GOLDEN, "test_it1_inversefisher_matrix.dat"))


def chunk_matrices():
    d = np.load(os.path.join(GOLDEN, "chunk_matrices.npz"))
    return d["signal_matrix"], d["q0_matrix"]


def all_close(a, b, rel=1e-5, abs_=1e-8):
    """isClose of tests/test_utils.cpp:7-11 and compareTestResults.py:8-9."""
    a, b = np.asarray(a), np.asarray(b)
    mag = np.maximum(np.abs(a), np.abs(b))
    return bool(np.all(np.abs(a - b) < abs_ + rel * mag))
