"""Copies the reference test suite's golden DATA into tests/golden/.

Run once in a container where the reference is mounted:
    python tests/golden/make_fixtures.py /root/reference

Only inputs and expected outputs are taken (the reference's tests/input and
tests/truth data files, plus the numeric truth arrays that
tests/cblas_tests.cpp:24-62,303-530 holds inline).  Large text matrices are parsed and
stored as compressed float64 arrays with the same values.
"""
import os
import re
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def read_text_matrix(path):
    with open(path) as f:
        nr, nc = (int(x) for x in f.readline().split())
        vals = np.array(f.read().split(), dtype=np.float64)
    return vals.reshape(nr, nc)


def inline_array(src, name):
    m = re.search(name + r"\[\]\s*=\s*\{([^}]*)\}", src)
    return np.array([float(x) for x in m.group(1).replace("\n", " ").split(",") if x.strip()])


def main(ref):
    t = os.path.join(ref, "tests")
    for rel in ["input/desilite_seed332298_id0_0_z2.4_lognormal.dat",
"input/slist.txt", "input/flist.txt",
                "truth/signal_R70000_dv20.0.fits",
                "truth/test_it1_fisher_matrix.dat",
                "truth/test_it1_inversefisher_matrix.dat",
                "truth/test_it1_quadratic_power_estimate_detailed.dat"]:
        shutil.copyfile(os.path.join(t, rel), os.path.join(HERE, os.path.basename(rel)))
    np.savez_compressed(
        os.path.join(HERE, "chunk_matrices.npz"),
        signal_matrix=read_text_matrix(os.path.join(t, "truth/signal_matrix.txt")),
        q0_matrix=read_text_matrix(os.path.join(t, "truth/q0_matrix.txt")))
    np.savez_compressed(
        os.path.join(HERE, "lu_inverse.npz"),
        matrix=read_text_matrix(os.path.join(t, "input/test_symmatrix.txt")),
        inverse=read_text_matrix(os.path.join(t, "truth/truth_inverse_matrix.txt")))
    with open(os.path.join(t, "cblas_tests.cpp")) as f:
        src = f.read()
    np.savez(os.path.join(HERE, "cblas_inline.npz"),
             sym_matrix_A=inline_array(src, "sym_matrix_A"),
             truth_lu_inversion_A=inline_array(src, "truth_lu_inversion_A"),
             matrix_for_SVD_A=inline_array(src, "matrix_for_SVD_A"),
             truth_SVD_A=inline_array(src, "truth_SVD_A"))
    # resolution-matrix storage (tests/cblas_tests.cpp:303-530)
    np.savez(os.path.join(HERE, "resolution_inline.npz"),
             **{name: inline_array(src, name) for name in (
                 "oversamp_values", "oversample_multiplier_A",
                 "truth_oversample_left_multiplication", "truth_oversample_right_multiplication",
                 "diamatrix_diagonals", "diamatrix_rows", "diamatrix_multiplier_B",
                 "truth_diamatrix_LN_multiplication", "truth_diamatrix_LT_multiplication",
                 "truth_diamatrix_RN_multiplication", "truth_diamatrix_RT_multiplication")})


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
