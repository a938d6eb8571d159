"""The FastBootstrap restatement (oracle/bootstrap.py) and the library's
Poisson draws (qml_poisson_draws, host code), on the CPU."""
import os

import numpy as np
import pytest

from lyspeq_amd import qml
from oracle import bootstrap as ob


def _draws(seed, n):
    if not os.path.exists(qml.LIB_PATH):
        pytest.skip("library not built (run `python -m lyspeq_amd.build`)")
    return qml.poisson_draws(seed, n)


def test_poisson_draws_are_unit_poisson_and_reproducible():
    n = 400000
    p = _draws(0, n)
    assert p.dtype == np.int32 and p.min() >= 0
    # mean and variance 1; P(0) = P(1) = 1/e, P(2) = 1/(2e); 5-sigma bands
    assert abs(p.mean() - 1.0) < 5 / np.sqrt(n)
    assert abs(p.var() - 1.0) < 5 * np.sqrt(3.0 / n)
    for k, pk in ((0, np.exp(-1)), (1, np.exp(-1)), (2, 0.5 * np.exp(-1))):
        frac = np.mean(p == k)
        assert abs(frac - pk) < 5 * np.sqrt(pk * (1 - pk) / n)
    assert np.array_equal(p[:1000], _draws(0, 1000))     # one generator, drawn in sequence
    assert not np.array_equal(p[:1000], _draws(1, 1000))


def _random_spectra(rng, nqso, T, nk):
    spectra = []
    for _ in range(nqso):
        chunks = []
        for _ in range(rng.integers(1, 4)):
            fis = int(rng.integers(0, T - 2 * nk))
            nq = int(rng.integers(nk, 2 * nk + 1))
            chunks.append((fis, nq, rng.standard_normal((3, nq))))
        spectra.append(chunks)
    return spectra


def test_unit_weights_give_the_summed_estimates_and_no_scatter():
    rng = np.random.default_rng(4)
    T, nqso, nboots = 60, 7, 9
    spectra = _random_spectra(rng, nqso, T, 10)
    col = [ob.collapse(s) for s in spectra]
    R = ob.realizations(np.ones((nqso, nboots), dtype=np.int32), col, T)
    total = np.zeros(T)
    for s in spectra:
        for fis, nq, dbt in s:
            total[fis:fis + nq] += dbt[0] - dbt[1] - dbt[2]
    assert np.allclose(R, total[None, :], rtol=1e-13, atol=1e-13)
    mean, cov = ob.mean_covariance(R)
    assert np.allclose(mean, total, rtol=1e-13, atol=1e-13)
    assert np.max(np.abs(cov)) <= 1e-28 * np.max(total ** 2)


def test_poisson_weights_give_the_sum_of_outer_products():
    """Independent unit-variance weights: Cov(R) = sum_q theta_q theta_q^T."""
    rng = np.random.default_rng(5)
    T, nqso, nboots = 40, 6, 40000
    spectra = _random_spectra(rng, nqso, T, 8)
    col = [ob.collapse(s) for s in spectra]
    P = _draws(3, nqso * nboots).reshape(nqso, nboots)
    R = ob.realizations(P, col, T)
    mean, cov = ob.mean_covariance(R)
    expect = np.zeros((T, T))
    emean = np.zeros(T)
    for istart, th in col:
        full = np.zeros(T)
        full[istart:istart + th.size] = th
        expect += np.outer(full, full)
        emean += full
    scale = np.sqrt(np.outer(np.diag(expect), np.diag(expect)))
    m = scale > 0
    # sampling error of a covariance entry: at most sqrt(2 / nboots) of the
    # scale for Gaussian weights, a little more for Poisson(1) (kurtosis 4)
    assert np.max(np.abs(cov - expect)[m] / scale[m]) < 8 * np.sqrt(2.0 / nboots)
    d = np.diag(expect) > 0
    assert np.max(np.abs(mean - emean)[d] / np.sqrt(np.diag(expect)[d])) < 6 / np.sqrt(nboots)


def test_brute_force_element_loop():
    rng = np.random.default_rng(6)
    T, nqso, nboots = 12, 3, 5
    spectra = _random_spectra(rng, nqso, T, 2)
    col = [ob.collapse(s) for s in spectra]
    P = rng.integers(0, 4, size=(nqso, nboots)).astype(np.int32)
    R = ob.realizations(P, col, T)
    for b in range(nboots):
        for i in range(T):
            s = 0.0
            for q, chunks in enumerate(spectra):
                th = 0.0
                for fis, nq, dbt in chunks:
                    if fis <= i < fis + nq:
                        th += dbt[0][i - fis] - dbt[1][i - fis] - dbt[2][i - fis]
                s += P[q, b] * th
            assert abs(R[b, i] - s) <= 1e-12 * max(1.0, abs(s))
    mean, cov = ob.mean_covariance(R)
    for i in range(T):
        for j in range(T):
            c = sum((R[b, i] - mean[i]) * (R[b, j] - mean[j]) for b in range(nboots)) / (nboots - 1)
            assert abs(cov[i, j] - c) <= 1e-12 * max(1.0, abs(c))


def _spd_chunks(rng, nqso, T, nk):
    spectra = []
    for _ in range(nqso):
        chunks = []
        for _ in range(rng.integers(1, 3)):
            fis = int(rng.integers(0, T - 2 * nk))
            nq = int(rng.integers(nk, 2 * nk + 1))
            a = rng.standard_normal((nq, nq + 4))
            chunks.append((fis, nq, np.triu(a @ a.T + nq * np.eye(nq)), rng.standard_normal((3, nq))))
        spectra.append(chunks)
    return spectra


def test_complete_form_with_unit_weights_is_the_power_estimate():
    """Every weight 1: F_b is the whole Fisher sum and x_b the driver's
    theta = 0.5 (0.5 F)^-1 (d - b - t) (oracle/driver.py, its own LU path)."""
    from oracle.driver import power_estimates
    rng = np.random.default_rng(9)
    T, nqso = 50, 8
    spectra = _spd_chunks(rng, nqso, T, 6)
    col = [ob.collapse([(f, n, dbt) for f, n, _, dbt in s]) for s in spectra]
    fis = [ob.collapse_fisher([(f, n, fc) for f, n, fc, _ in s]) for s in spectra]
    X, valid = ob.complete_realizations(np.ones((3, nqso), dtype=np.int32), col, fis, T)
    assert valid.all() and X.shape == (3, T)
    Fsum, dsum = np.zeros((T, T)), np.zeros((3, T))
    for s in spectra:
        for f, n, fc, dbt in s:
            Fsum[f:f + n, f:f + n] += fc
            dsum[:, f:f + n] += dbt
    res = power_estimates(Fsum, dsum)
    assert res["damp"] == 0
    for x in X:
        assert np.max(np.abs(x - res["theta"])) <= 1e-10 * np.max(np.abs(res["theta"]))


def test_complete_form_drops_realisations_that_do_not_factorise():
    rng = np.random.default_rng(10)
    T, nqso = 30, 4
    spectra = _spd_chunks(rng, nqso, T, 5)
    col = [ob.collapse([(f, n, dbt) for f, n, _, dbt in s]) for s in spectra]
    fis = [ob.collapse_fisher([(f, n, fc) for f, n, fc, _ in s]) for s in spectra]
    W = np.ones((4, nqso), dtype=np.int32)
    W[2, 0] = -3          # a negative weight makes F_b indefinite
    X, valid = ob.complete_realizations(W, col, fis, T)
    assert list(valid) == [True, True, False, True] and X.shape == (3, T)
    W[1] = 0              # no spectrum: every band empty, x = 0
    X, valid = ob.complete_realizations(W, col, fis, T)
    assert valid[1] and not X[1].any()


def _brute_median(vals):
    s = sorted(vals)
    n = len(s)
    return s[n // 2] if n % 2 else (s[n // 2 - 1] + s[n // 2]) / 2


@pytest.mark.parametrize("nboots", [7, 8])
def test_median_bootstrap_brute_force(nboots):
    """Sorted Python lists, band by band and pair by pair, as
    _calcuateMedian and _calcuateMadCovariance state it."""
    rng = np.random.default_rng(nboots)
    R = np.round(rng.standard_normal((nboots, 4)), 1)        # ties included
    med, mad = ob.median_covariance(R)
    T = R.shape[1]
    bm = [_brute_median([R[b][i] for b in range(nboots)]) for i in range(T)]
    dev = [[R[b][i] - bm[i] for b in range(nboots)] for i in range(T)]
    assert list(med) == bm
    for i in range(T):
        for j in range(T):
            lo, hi = min(i, j), max(i, j)
            assert mad[i, j] == _brute_median([dev[lo][b] * dev[hi][b] for b in range(nboots)])


@pytest.mark.parametrize("nboots", [1, 2, 999, 1000])
def test_median_is_numpys_and_the_odd_diagonal_is_the_squared_mad(nboots):
    from scipy.stats import median_abs_deviation
    rng = np.random.default_rng(nboots + 5)
    R = rng.standard_normal((nboots, 6)) * np.arange(1, 7) + 3.0
    med, mad = ob.median_covariance(R)
    assert np.array_equal(med, np.median(R, axis=0))
    assert np.array_equal(mad, mad.T)
    if nboots % 2:
        np.testing.assert_allclose(np.diag(mad), median_abs_deviation(R, axis=0, scale=1.0) ** 2, rtol=1e-15, atol=0)
    # the median covariance does not depend on the order of the realisations
    perm = rng.permutation(nboots)
    m2, c2 = ob.median_covariance(R[perm])
    assert np.array_equal(m2, med) and np.array_equal(c2, mad)
