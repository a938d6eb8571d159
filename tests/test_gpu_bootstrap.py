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
