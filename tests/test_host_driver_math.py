"""The C++ Newton-Raphson driver arithmetic (lyspeq_amd/host/qml_driver.cpp)
against the reference's golden iteration files, its LU-inverse truth, numpy's
singular values, closed forms and the oracle's restatement.  Host code only:
runs on the CPU through the compiled checker tests/cpp/driver_check."""
import os
import struct
import subprocess

import numpy as np
import pytest

import goldencase as gc
from lyspeq_amd import build as B
from oracle.driver import has_converged, power_estimates


def _call(tmp_path, cmd, ints, arrays):
    if not os.path.exists(B.DRIVER_CHECK):
        pytest.skip("tests/cpp/driver_check not built (run `python -m lyspeq_amd.build`)")
    inp, out = tmp_path / f"{cmd}.in", tmp_path / f"{cmd}.out"
    with open(inp, "wb") as f:
        for i in ints:
            f.write(struct.pack("<i", int(i)))
        for a in arrays:
            f.write(np.ascontiguousarray(a, "<f8").tobytes())
    r = subprocess.run([B.DRIVER_CHECK, cmd, str(inp), str(out)], capture_output=True, text=True, timeout=120)
    return r, (np.fromfile(out, "<f8") if r.returncode == 0 else None)


def _power(tmp_path, fisher_sum, dbt):
    T = fisher_sum.shape[0]
    r, o = _call(tmp_path, "power", [T], [fisher_sum, dbt])
    assert r.returncode == 0, r.stderr
    TT = T * T
    parts = dict(fisher=o[:TT].reshape(T, T), solver=o[TT:2 * TT].reshape(T, T),
                 inverse=o[2 * TT:3 * TT].reshape(T, T), theta=o[3 * TT:3 * TT + T],
                 sigma=o[3 * TT + T:3 * TT + 2 * T], weighted=o[3 * TT + 2 * T:3 * TT + 5 * T].reshape(3, T),
                 damp=o[3 * TT + 5 * T], dof=int(o[3 * TT + 5 * T + 1]))
    assert o.size == 3 * TT + 5 * T + 2
    return parts


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - b)) / np.max(np.abs(b)))


def test_lu_inverse_against_reference_truth(tmp_path, golden_dir):
    # tests/cblas_tests.cpp:257-283: LU inverse of the 496 x 496 test matrix
    d = np.load(f"{golden_dir}/lu_inverse.npz")
    A = d["matrix"]
    n = A.shape[0]
    r, inv = _call(tmp_path, "invert", [n], [A])
    assert r.returncode == 0, r.stderr
    inv = inv.reshape(n, n)
    assert gc.all_close(inv, d["inverse"])
    assert np.max(np.abs(A @ inv - np.eye(n))) < 1e-9


def test_singular_matrix_raises_the_reference_message(tmp_path):
    A = np.ones((4, 4))
    r, _ = _call(tmp_path, "invert", [4], [A])
    assert r.returncode == 3 and "ERROR in LU decomposition." in r.stderr


def test_singular_values_match_numpy(tmp_path):
    rng = np.random.default_rng(7)
    n = 60
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.concatenate([np.geomspace(1e3, 1e-3, n - 3), [-2.5, -40.0, 0.0]])
    A = (Q * lam) @ Q.T
    A = 0.5 * (A + A.T)
    r, sv = _call(tmp_path, "svals", [n], [A])
    assert r.returncode == 0, r.stderr
    ref = np.linalg.svd(A, compute_uv=False)
    assert np.all(np.diff(sv) <= 0)
    assert np.max(np.abs(sv - ref)) < 1e-11 * ref[0]
    # a diagonal matrix: the magnitudes of its diagonal, exactly up to ordering
    D = np.diag([3.0, -7.0, 0.5, 2.0])
    r, sv = _call(tmp_path, "svals", [4], [D])
    assert np.allclose(sv, [7.0, 3.0, 2.0, 0.5], rtol=1e-15, atol=0)


def test_power_estimates_against_golden_iteration(tmp_path):
    """tests/truth/test_it1_*: the halved Fisher sum, its inverse, theta,
    sigma and the Fisher-weighted d, b, t.  The input is the golden chunk's
    unhalved Fisher and raw d, b, t sums at full precision (the oracle's, whose
    own test pins them against the same files): the golden Fisher file keeps 7
    digits, too few to invert this ill-conditioned matrix to the golden inverse."""
    from oracle.bins import Bins
    from oracle.chunk import Chunk, TableSet
    b = Bins(**gc.BINS_KW)
    t = gc.table()
    sp = gc.spectrum()
    ch = Chunk(b, [TableSet(70000, 20.0, t.nvpoints, t.v_length, t.nzpoints, t.deriv, t.signal)],
               sp.wave, sp.flux, sp.ivar, sp.R_fwhm)
    T = b.total_kz_bins
    F, dbt = np.zeros((T, T)), np.zeros((3, T))
    assert ch.one_iteration(np.zeros(T), F, dbt)
    rows = gc.detailed_rows()
    for i in range(3):
        assert gc.all_close(dbt[i], rows[:, 11 + i])
    p = _power(tmp_path, F, dbt)
    assert gc.all_close(p["fisher"], gc.golden_fisher())
    assert gc.all_close(p["inverse"], gc.golden_inverse_fisher())
    assert gc.all_close(p["theta"], rows[:, 5])
    assert gc.all_close(p["sigma"], rows[:, 7])
    for i in range(3):
        assert gc.all_close(p["weighted"][i], rows[:, 8 + i])
    assert p["damp"] == 0.0 and p["dof"] == T
    o = power_estimates(F, dbt)
    assert _rel(p["theta"], o["theta"]) < 1e-10 and _rel(p["inverse"], o["inverse"]) < 1e-10


def test_damped_inversion_with_an_empty_diagonal_against_oracle(tmp_path):
    """A singular-value jump above 8 (1 -> 0.05 -> 0.001) sets the damping to
    0.05, and a band with a zero row and column is left out of dof."""
    rng = np.random.default_rng(3)
    n = 40
    Q, _ = np.linalg.qr(rng.standard_normal((n - 1, n - 1)))
    lam = np.concatenate([np.linspace(10.0, 1.0, n - 3), [0.05, 0.001]])
    F = (Q * lam) @ Q.T
    F = 0.5 * (F + F.T)
    F = np.insert(np.insert(F, 7, 0.0, axis=0), 7, 0.0, axis=1)
    dbt = rng.standard_normal((3, n))
    p = _power(tmp_path, 2.0 * np.triu(F), dbt)
    o = power_estimates(2.0 * np.triu(F), dbt)
    assert o["damped"] and abs(p["damp"] - 0.05) < 1e-12 and abs(o["damp"] - 0.05) < 1e-12
    assert p["dof"] == o["dof"] == n - 1
    for key in ("solver", "inverse", "theta", "weighted"):
        assert _rel(p[key], o[key]) < 1e-10, key
    assert p["inverse"][7, 7] == 0.0 and p["solver"][7, 7] == 0.0 and p["sigma"][7] == 0.0
    # the damped solver is (F F + damp I)^-1 F on the filled matrix
    Fd = F.copy()
    Fd[7, 7] = 1.0
    S = np.linalg.solve(Fd @ Fd + 0.05 * np.eye(n), Fd)
    S[7, 7] = 0.0
    assert _rel(p["solver"], S) < 1e-10


def test_convergence_test_closed_forms(tmp_path):
    """previous = 2 current gives relative changes of 1/2; steps of
    sqrt(e_i) against an inverse diag(e) give chi2_diag = sqrt(m / dof) over
    the m bands with e_i > 0; a Fisher c I gives chi2_full = sqrt(c sum t^2 / dof)."""
    T, dof = 6, 5
    e = np.array([1.0, 4.0, 9.0, 0.0, 2.0, 3.0])
    cur = np.array([1.0, 2.0, 3.0, 0.0, 0.5, 4.0])
    prev = 2.0 * cur
    prev[3] = 0.0
    inv = np.diag(e)
    t = prev - cur
    c = 0.25
    fis = c * np.eye(T)
    for eps, conv in ((10.0, True), (1e-3, False)):
        r, o = _call(tmp_path, "converge", [T, dof], [[eps], cur, prev, inv, fis])
        assert r.returncode == 0, r.stderr
        abs_mean, abs_max, chi2_diag, chi2_full, converged = o
        assert abs(abs_max - 0.5) < 1e-15 and abs(abs_mean - 0.5 * 5 / dof) < 1e-15
        ok = e > 0
        assert abs(chi2_diag - np.sqrt(np.sum(t[ok] ** 2 / e[ok]) / dof)) < 1e-14
        assert abs(chi2_full - np.sqrt(c * np.sum(t ** 2) / dof)) < 1e-14
        assert bool(converged) == conv
        ref = has_converged(cur, prev, inv, fis, dof, eps)
        assert np.allclose(o[:4], ref[:4], rtol=1e-14, atol=0) and bool(converged) == ref[4]
    # a step of one sigma per band: chi2_diag = 1 when every band counts
    e2 = np.array([1.0, 4.0, 9.0, 16.0, 2.0, 3.0])
    prev2 = cur + np.sqrt(e2)
    r, o = _call(tmp_path, "converge", [T, T], [[1.0 + 1e-12], cur, prev2, np.diag(e2), np.zeros((T, T))])
    assert abs(o[2] - 1.0) < 1e-14 and o[3] == 0.0 and o[4] == 1.0


def test_script_smoother_runs_the_smoothing_script_like_the_reference(tmp_path):
    """ScriptSmoother (the reference's _smoothPowerSpectra: the spectrum file
    of writeSpectrumEstimates, the script run as a child process, one value
    per band read back, P_fid subtracted) running python3 -m
    lyspeq_amd.smooth must give the Python driver's smooth_power fits bit
    for bit, in both modes; a failing command raises "smoothing error"."""
    import sys
    from lyspeq_amd.driver import smooth_power
    sys.path.insert(0, os.path.dirname(__file__))
    from test_driver_smoother import _grid, _surface
    kc, zc = _grid()
    truth = _surface(kc, zc)
    pfid = 0.9 * truth
    sigma = 0.05 * truth
    rng = np.random.default_rng(4)
    theta = truth + sigma * rng.standard_normal(truth.size) - pfid
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    cmd = f"PYTHONPATH={root} {sys.executable} -m lyspeq_amd.smooth"
    env = dict(os.environ, TMPDIR=str(tmp_path))
    for lg in (0, 1):
        inp, out = tmp_path / f"s{lg}.in", tmp_path / f"s{lg}.out"
        with open(inp, "wb") as f:
            f.write(np.array([zc.size, kc.size, lg], dtype="<i4").tobytes())
            for a in (zc, kc, theta, sigma, pfid):
                f.write(np.asarray(a, dtype="<f8").tobytes())
        r = subprocess.run([B.DRIVER_CHECK, "smooth", str(inp), str(out), cmd], capture_output=True, text=True,
                           timeout=120, env=env)
        assert r.returncode == 0, r.stderr
        got = np.fromfile(out, dtype="<f8")
        want = smooth_power(theta, sigma, kc, zc, pfid, interp_log=bool(lg))
        assert np.array_equal(got, want)
    assert not any(p.name.startswith("tmp-") for p in tmp_path.iterdir())   # the files are removed
    r = subprocess.run([B.DRIVER_CHECK, "smooth", str(inp), str(out), "false"], capture_output=True, text=True,
                       timeout=60, env=env)
    assert r.returncode == 3 and "smoothing error" in r.stderr


def test_power_estimates_at_the_desi_band_count_against_oracle(tmp_path):
    """T = 455 (35 k bands x 13 z bins) as the DESI-like workload has it:
    the Fisher sum of chunks that each cover up to 2 z bins and a band range
    that ends at their KBIN_UPP, so high-k bands of some z bins are empty and
    the matrix is banded by z bin.  The C++ arithmetic (normalise, stable
    inversion with the empty-diagonal rule and the singular-value test,
    theta, sigma, the weighted d/b/t) against the oracle's restatement."""
    rng = np.random.default_rng(455)
    nk, nz = 35, 13
    T = nk * nz
    F = np.zeros((T, T))
    dbt = np.zeros((3, T))
    for _ in range(600):
        z0 = int(rng.integers(0, nz - 1))
        kup = int(rng.integers(10, nk + 1)) if z0 < nz - 3 else int(rng.integers(5, 12))
        idx = np.concatenate([np.arange(z0 * nk, z0 * nk + kup), np.arange((z0 + 1) * nk, (z0 + 1) * nk + kup)])
        A = rng.standard_normal((idx.size, 3 * idx.size)) / idx.size
        blk = A @ A.T
        F[np.ix_(idx, idx)] += blk
        dbt[:, idx] += rng.standard_normal((3, idx.size)) * np.sqrt(np.diag(blk))
    F = 2.0 * np.triu(F)   # the unhalved upper triangle the chunks accumulate
    empty = int(np.sum(np.diag(F) == 0))
    assert empty > 0
    p = _power(tmp_path, F, dbt)
    o = power_estimates(F, dbt)
    assert p["dof"] == o["dof"] == T - empty
    assert abs(p["damp"] - o["damp"]) <= 1e-12 * max(1.0, abs(o["damp"]))
    for key in ("fisher", "solver", "inverse", "theta", "weighted"):
        assert _rel(p[key], o[key]) < 1e-10, key
    ok = o["sigma"] > 0
    assert np.all(np.abs(p["sigma"][ok] - o["sigma"][ok]) <= 1e-10 * o["sigma"][ok])
