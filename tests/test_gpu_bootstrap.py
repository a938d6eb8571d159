"""FastBootstrap on the device (qml_bootstrap) against the oracle's
restatement (oracle/bootstrap.py) on the same per-chunk results and the same
Poisson draws (needs an MI355X)."""
import numpy as np
import pytest

import parity as P
from lyspeq_amd import qml, synth
from lyspeq_amd.qml import QMLEngine
from oracle import bootstrap as ob

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_fast_bootstrap_matches_oracle(precision):
    print(f"[case] fast bootstrap {precision}", flush=True)
    w = synth.desi_config(nchunks=24)
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 12)
    eng = QMLEngine(w.binning, precision=precision, max_batch_chunks=9)
    accepted = synth.load_engine(eng, w, list(range(len(w.chunks))))
    eng.commit()
    T = eng.T
    F, dbt = np.zeros((T, T)), np.zeros((3, T))
    eng.iterate(th, F, dbt)
    nch = len(eng.chunk_infos)
    # spectra of 1-4 consecutive chunks in queue order; chunk 5 left out
    rng = np.random.default_rng(2)
    qso = np.full(nch, -1, dtype=np.int32)
    q, i = 0, 0
    while i < nch:
        k = int(rng.integers(1, 5))
        qso[i:i + k] = q
        i += k
        q += 1
    nqso = q
    qso[5] = -1
    nboots, seed = 3000, 7
    R, mean, cov = eng.bootstrap(nboots, qso, nqso, seed=seed)
    # oracle on the engine's own per-chunk results and the same draws
    res = [eng.chunk_result(c) for c in range(nch)]
    spectra = [[(eng.chunk_infos[c].fisher_index_start, eng.chunk_infos[c].n_q_matrices, res[c][1])
                for c in range(nch) if qso[c] == s] for s in range(nqso)]
    assert all(spectra)
    col = [ob.collapse(s) for s in spectra]
    Pw = qml.poisson_draws(seed, nqso * nboots).reshape(nqso, nboots)
    Ro = ob.realizations(Pw, col, T)
    mo, co = ob.mean_covariance(Ro)
    er, em, ec = _rel(R, Ro), _rel(mean, mo), _rel(cov, co)
    print(f"bootstrap {precision}: {nqso} spectra, {nboots} realisations; rel err R {er:.1e} "
          f"mean {em:.1e} cov {ec:.1e}", flush=True)
    assert er < 1e-12 and em < 1e-12 and ec < 1e-11
    # unit weights would give the accumulated d - b - t of the included
    # chunks: the mean of Poisson(1)-weighted realisations scatters about it
    inc = np.zeros(T)
    for c in range(nch):
        if qso[c] >= 0:
            info = eng.chunk_infos[c]
            f, n = info.fisher_index_start, info.n_q_matrices
            inc[f:f + n] += res[c][1][0] - res[c][1][1] - res[c][1][2]
    sd = np.sqrt(np.diag(co))
    ok = sd > 0
    assert np.max(np.abs(mean - inc)[ok] / sd[ok]) < 6 / np.sqrt(nboots)
    # every chunk in the sums: the collapsed vectors add up to the engine's
    # accumulated d - b - t
    full = np.zeros(nch, dtype=np.int32)
    _, m1, _ = eng.bootstrap(2, full, 1, seed=seed, realizations=False)
    p1 = qml.poisson_draws(seed, 2).astype(np.float64)
    tot = dbt[0] - dbt[1] - dbt[2]
    # (d, b and t cancel in d - b - t: the tolerance is set by their size)
    assert np.allclose(m1, tot * p1.mean(), rtol=0, atol=1e-12 * np.abs(dbt).max() * max(1.0, p1.max()))
    eng.close()


def test_bootstrap_arguments_are_checked():
    print("[case] bootstrap arguments", flush=True)
    w = synth.desi_config(nchunks=3)
    eng = QMLEngine(w.binning, precision="fp64")
    synth.load_engine(eng, w, list(range(len(w.chunks))))
    eng.commit()
    nch = len(eng.chunk_infos)
    with pytest.raises(Exception):
        eng.bootstrap(1, np.zeros(nch), 1)                 # one realisation has no covariance
    with pytest.raises(Exception):
        eng.bootstrap(10, np.full(nch, 3), 2)              # spectrum index beyond nqso
    eng.close()


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_complete_bootstrap_matches_oracle(precision):
    """The complete form on the device (a Cholesky solve per realisation)
    against the oracle's, on the same per-chunk results and draws; with unit
    weights every realisation is the solution of the whole Fisher sum."""
    print(f"[case] complete bootstrap {precision}", flush=True)
    w = synth.desi_config(nchunks=24)
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 14)
    eng = QMLEngine(w.binning, precision=precision, max_batch_chunks=9)
    synth.load_engine(eng, w, list(range(len(w.chunks))))
    eng.commit()
    T = eng.T
    F, dbt = np.zeros((T, T)), np.zeros((3, T))
    eng.iterate(th, F, dbt)
    nch = len(eng.chunk_infos)
    qso = (np.arange(nch) // 3).astype(np.int32)       # spectra of three chunks
    nqso = int(qso.max()) + 1
    res = [eng.chunk_result(c) for c in range(nch)]
    infos = eng.chunk_infos
    col = [ob.collapse([(infos[c].fisher_index_start, infos[c].n_q_matrices, res[c][1])
                        for c in range(nch) if qso[c] == s]) for s in range(nqso)]
    fis = [ob.collapse_fisher([(infos[c].fisher_index_start, infos[c].n_q_matrices, res[c][0])
                               for c in range(nch) if qso[c] == s]) for s in range(nqso)]
    nboots, seed = 100, 5
    X, valid, mean, cov = eng.bootstrap_complete(nboots, qso, nqso, seed=seed)
    Wd = qml.poisson_draws(seed, nboots * nqso).reshape(nboots, nqso)
    Xo, vo, systems = ob.complete_realizations(Wd, col, fis, T, return_systems=True)
    assert np.array_equal(valid, vo) and X.shape == Xo.shape
    # two correct Cholesky solves agree to about n cond eps, with the
    # condition number of the diagonally equilibrated F_b (Cholesky is
    # insensitive to diagonal scaling), per band relative to the band's own
    # scale; the device solves through the explicit triangular inverse
    eps = np.finfo(float).eps

    def _kappa(S):
        d = np.sqrt(np.diag(S))
        return np.linalg.cond(S / np.outer(d, d))

    worst, emax, kmax = 0.0, 0.0, 0.0
    for xb, xo, (S, r) in zip(X, Xo, systems):
        kappa = _kappa(S)
        kmax = max(kmax, kappa)
        d = np.sqrt(np.diag(S))
        e = float(np.max(np.abs((xb - xo) * d)) / np.max(np.abs(xo * d)))
        emax = max(emax, e)
        bar = 20 * T * kappa * eps
        worst = max(worst, e / bar)
        assert e < bar
    # mean and covariance of the device's realisations
    mg, cg = ob.mean_covariance(X)
    em, ec = _rel(mean, mg), _rel(cov, cg)
    print(f"complete bootstrap {precision}: {valid.sum()} of {nboots} valid; equilibrated cond(F_b) up to "
          f"{kmax:.1e}; realisations differ by {emax:.1e} (scaled), {worst:.3f} of the 20 n cond eps bar; "
          f"mean {em:.1e} cov {ec:.1e}", flush=True)
    assert em < 1e-12 and ec < 1e-11
    # unit weights: every realisation is (sum of the chunk Fisher matrices)^-1 (d - b - t)
    X1, v1, _, _ = eng.bootstrap_complete(4, qso, nqso, weights=np.ones((4, nqso), dtype=np.int32))
    S = np.triu(F) + np.triu(F, 1).T
    empty = np.diag(S) == 0
    S[empty, empty] = 1.0
    x = np.linalg.solve(S, dbt[0] - dbt[1] - dbt[2])
    x[empty] = 0.0
    assert v1.all()
    d = np.sqrt(np.diag(S))
    tol = 20 * T * _kappa(S) * eps
    for xb in X1:
        assert np.max(np.abs((xb - x) * d)) / np.max(np.abs(x * d)) < tol
    eng.close()


def _median_engine():
    w = synth.cpu_reference_config(nchunks=2)
    return QMLEngine(w.binning, precision="fp64")


@pytest.mark.parametrize("nboots,T,decimals", [(1, 3, None), (2, 5, None), (2001, 37, None), (2000, 70, 1),
                                               (1000, 130, 1), (20000, 12, 0), (19999, 9, 0)])
def test_median_bootstrap_matches_oracle(nboots, T, decimals):
    """qml_bootstrap_median against the oracle bit for bit: the radix
    select returns the sorted arrays' values, and the half-sums and products
    are the same FP64 operations.  Rounded inputs give ties and zero
    deviations (signed zeros among the products).  Integer-rounded inputs
    put thousands of equal values in the median's bin, beyond what the
    kernel keeps in LDS: those selects read the values in every pass."""
    print(f"[case] median bootstrap nboots={nboots} T={T} decimals={decimals}", flush=True)
    rng = np.random.default_rng(nboots + T)
    if decimals == 0:
        R = np.round(rng.standard_normal((nboots, T)) * rng.uniform(0.8, 1.5, T))
    else:
        R = rng.standard_normal((nboots, T)) * rng.uniform(0.1, 10.0, T) + rng.uniform(-1.0, 1.0, T)
    if decimals:
        R = np.round(R, decimals)
    eng = _median_engine()
    med, mad = eng.bootstrap_median(R)
    eng.close()
    omed, omad = ob.median_covariance(R)
    assert np.array_equal(med, omed)
    assert np.array_equal(mad, omad)


def test_median_bootstrap_full_size():
    """20000 realisations of 455 bands (the DESI binning): every median, and
    3000 pairs of the median covariance against the oracle, bit for bit."""
    import time
    nboots, T = 20000, 455
    print(f"[case] median bootstrap full size nboots={nboots} T={T}", flush=True)
    rng = np.random.default_rng(455)
    R = rng.standard_normal((nboots, T)) @ (np.eye(T) + 0.1 * rng.standard_normal((T, T)) / np.sqrt(T))
    eng = _median_engine()
    eng.bootstrap_median(R[:64, :16])                       # load the kernels
    t0 = time.perf_counter()
    med, mad = eng.bootstrap_median(R)
    ms = (time.perf_counter() - t0) * 1e3
    eng.close()
    print(f"median bootstrap {nboots} x {T}: {ms:.1f} ms with transfers", flush=True)
    assert np.array_equal(med, ob.median(R))
    assert np.array_equal(mad, mad.T)
    X = R.T - med[:, None]
    ii = rng.integers(0, T, 3000)
    jj = rng.integers(0, T, 3000)
    for s in range(0, 3000, 500):
        i, j = ii[s:s + 500], jj[s:s + 500]
        want = ob._median_rows(X[i] * X[j])
        assert np.array_equal(mad[i, j], want)
