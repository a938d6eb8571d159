"""Edge cases of the chunk step (needs an MI355X): the constructor's
rejections against the oracle's, masked boundaries, call-order errors, and
a covariance the factorisation rejects."""
import numpy as np
import pytest

import parity as P
from lyspeq_amd import synth
from lyspeq_amd.qml import EmptyChunkError, QMLEngine, QMLError, ShortChunkError, TableNotFoundError
from oracle.chunk import Chunk as OracleChunk, ChunkError

pytestmark = pytest.mark.gpu



def _print_case(name):
    print(f"[case] {name}", flush=True)


def _engine(w, precision="fp64", **kw):
    eng = QMLEngine(w.binning, precision=precision, **kw)
    for i, (R, dv) in enumerate(w.pairs):
        eng.add_table(R, dv, w.deriv[i], w.signal[i], w.v_length)
    return eng


def test_constructor_rejections_match_the_oracle():
    """Short chunk (core/chunk_estimate.cpp:172-187), no table for (R, dv)
    (:198-206) and all pixels masked (io/qso_file.cpp:150-153): the engine
    refuses the chunk with the matching error where the oracle's constructor
    raises; a chunk whose masked boundary is cut keeps its inner pixels."""
    _print_case("constructor rejections")
    w = synth.cpu_reference_config(nchunks=2)
    b, tabs = P.oracle_bins(w.binning), P.oracle_tables(w)
    c = w.chunks[0]
    wave, delta, ivar, R = c["wave"], c["delta"], c["ivar"], c["R"]
    eng = _engine(w)
    cases = [
        ("short", dict(wave=wave[:100], delta=delta[:100], ivar=ivar[:100], R=R), ShortChunkError, "Short"),
        ("no table", dict(wave=wave, delta=delta, ivar=ivar, R=R + 7), TableNotFoundError, "not found"),
        ("masked", dict(wave=wave, delta=delta, ivar=np.zeros_like(ivar), R=R), EmptyChunkError, "Empty"),
    ]
    for name, kw, err, words in cases:
        with pytest.raises(ChunkError, match=words):
            OracleChunk(b, tabs, kw["wave"], kw["delta"], kw["ivar"], kw["R"])
        with pytest.raises(err):
            eng.add_chunk(kw["wave"], kw["delta"], kw["ivar"], kw["R"])
        print(f"{name}: rejected by both", flush=True)
    # masked boundary: 5 pixels at the start and 7 at the end are cut
    iv = ivar.copy()
    iv[:5] = 0.0
    iv[-7:] = 0.0
    cid, info = eng.add_chunk(wave, delta, iv, R)
    assert info.size == wave.size - 12
    eng.commit()
    T = eng.T
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 21)
    F, d = np.zeros((T, T)), np.zeros((3, T))
    eng.iterate(th, F, d)
    Fc, dbt, st = eng.chunk_result(cid)
    eng.close()
    och = OracleChunk(b, tabs, wave, delta, iv, R)
    assert och.N == info.size
    oF, od = np.zeros((T, T)), np.zeros((3, T))
    assert och.one_iteration(th, oF, od)
    assert st == 0
    P.assert_fp64(F, d, oF, od, "masked boundary cut")


def test_call_order_and_argument_errors():
    _print_case("call order")
    w = synth.cpu_reference_config(nchunks=2)
    eng = _engine(w)
    T = w.binning.total_kz_bins
    F, d = np.zeros((T, T)), np.zeros((3, T))
    with pytest.raises(QMLError):
        eng.iterate(np.zeros(T), F, d)               # before commit
    c = w.chunks[0]
    cid, _ = eng.add_chunk(c["wave"], c["delta"], c["ivar"], c["R"])
    eng.commit()
    c = w.chunks[1]
    with pytest.raises(QMLError):
        eng.add_chunk(c["wave"], c["delta"], c["ivar"], c["R"])   # after commit
    with pytest.raises(QMLError):
        eng.commit()                                 # twice
    with pytest.raises(QMLError):
        eng.set_chunk_active(cid + 5, False)         # unknown chunk
    eng.iterate(np.zeros(T), F, d)
    assert np.abs(F).max() > 0
    eng.close()


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_rejected_factorisation_leaves_the_chunk_out(precision):
    """theta = -s P_fid makes C = S + sum theta Q indefinite for large s: the
    Cholesky factorisation of I + D C D fails, the chunk reports status 1,
    keeps zero results and is left out of the sums, the counterpart of the
    reference's caught LAPACK failure (core/chunk_estimate.cpp:905-913).  At
    every s the sums are the sums over the chunks that passed."""
    _print_case(f"rejected factorisation {precision}")
    w = synth.cpu_reference_config(nchunks=8)
    eng = _engine(w, precision)
    acc = []
    for i, c in enumerate(w.chunks):
        cid, info = eng.add_chunk(c["wave"], c["delta"], c["ivar"], c["R"])
        acc.append((cid, info))
    eng.commit()
    T = eng.T
    kc = w.binning.kband_centers()
    zc = w.binning.zbin_centers()
    pfid = np.array([[synth.pd13_power(k, z, *w.pd13) for k in kc] for z in zc]).reshape(-1)
    seen = set()
    for s in (0.0, 0.5, 2.0, 30.0):
        th = -s * pfid
        F, d = np.zeros((T, T)), np.zeros((3, T))
        eng.iterate(th, F, d)
        Fs, ds = np.zeros((T, T)), np.zeros((3, T))
        nfail = 0
        for cid, info in acc:
            Fc, dbt, st = eng.chunk_result(cid)
            if st != 0:
                nfail += 1
                assert st == 1 and not Fc.any() and not dbt.any()
                continue
            f, n = info.fisher_index_start, info.n_q_matrices
            Fs[f:f + n, f:f + n] += np.triu(Fc)
            ds[:, f:f + n] += dbt
        assert np.allclose(np.triu(F), Fs, rtol=1e-13, atol=0) and np.allclose(d, ds, rtol=1e-13, atol=0)
        seen.add(nfail)
        print(f"theta = -{s} P_fid: {nfail} of {len(acc)} chunks rejected", flush=True)
        if s == 0.0:
            assert nfail == 0
        if s == 30.0:
            assert nfail == len(acc) and not F.any() and not d.any()
    eng.close()


def test_chunk_beyond_the_scratch_budget_is_refused_at_registration():
    """Chunk::_setStoredMatrices (core/chunk_estimate.cpp:258-282) throws
    "Not all matrices are stored." and the chunk is skipped; here a chunk
    whose scratch exceeds the engine's budget is refused by qml_add_chunk,
    and a smaller one is committed and iterated as usual.  Under a 0.1 GB
    budget a DESI chunk of N >= 500 needs about 0.25 GB or more in split
    BF16, one of N <= 250 a few tens of MB."""
    from lyspeq_amd.qml import ChunkMemoryError
    _print_case("scratch budget")
    wd = synth.desi_config(nchunks=24)
    by_size = sorted(wd.chunks, key=lambda c: len(c["wave"]))
    c = by_size[-1]
    assert len(c["wave"]) >= 500
    eng = QMLEngine(wd.binning, precision="bf16x3", scratch_gb=0.1)
    for i, (R, dv) in enumerate(wd.pairs):
        eng.add_table(R, dv, wd.deriv[i], wd.signal[i], wd.v_length)
    with pytest.raises(ChunkMemoryError, match="Not all matrices are stored"):
        eng.add_chunk(c["wave"], c["delta"], c["ivar"], c["R"])
    cid = None
    for small in by_size[:4]:          # the shortest chunk the constructor accepts
        try:
            cid, info = eng.add_chunk(small["wave"], small["delta"], small["ivar"], small["R"])
            break
        except (ShortChunkError, TableNotFoundError, EmptyChunkError):
            continue
    assert cid is not None and info.size <= 250
    eng.commit()
    T = eng.T
    F, d = np.zeros((T, T)), np.zeros((3, T))
    eng.iterate(np.zeros(T), F, d)
    assert np.abs(F).max() > 0 and eng.chunk_result(cid)[2] == 0 and cid == 0
    eng.close()


def test_redshift_bin_limit_of_a_chunk():
    """A chunk may span at most 8 redshift bins in this build (the kernels
    pack a chunk's z-bin weights for 8 bins).  With bins of width 0.03 the
    CPU configuration's chunk spans exactly 8 (triangular binning) and must
    meet the FP64 bar against the oracle; with width 0.015 it spans 14 and
    registration refuses it with QML_ERR_LIMIT, loudly: the reference would
    process it, so the chunk is not skipped in silence."""
    import dataclasses
    _print_case("z-bin limit")
    w = synth.cpu_reference_config(nchunks=2)
    c = w.chunks[0]
    tabs = P.oracle_tables(w)
    for width, nz, spans in ((0.03, 9, 8), (0.015, 17, 14)):
        b = dataclasses.replace(w.binning, z0=2.30, z_bin_width=width, nz=nz)
        ob = P.oracle_bins(b)
        och = OracleChunk(ob, tabs, c["wave"], c["delta"], c["ivar"], c["R"])
        assert och.n_q == spans * ob.nk   # N_Q_MATRICES: the z bins spanned times the k bands
        eng = QMLEngine(b, precision="fp64")
        for i, (R, dv) in enumerate(w.pairs):
            eng.add_table(R, dv, w.deriv[i], w.signal[i], w.v_length)
        if spans > 8:
            with pytest.raises(QMLError) as e:
                eng.add_chunk(c["wave"], c["delta"], c["ivar"], c["R"])
            assert e.value.code == -7 and "more than 8 redshift bins" in str(e.value)
            eng.close()
            continue
        cid, info = eng.add_chunk(c["wave"], c["delta"], c["ivar"], c["R"])
        eng.commit()
        T = eng.T
        th = np.zeros(T)
        F, d = np.zeros((T, T)), np.zeros((3, T))
        eng.iterate(th, F, d)
        eng.close()
        Fo, do = np.zeros((T, T)), np.zeros((3, T))
        assert och.one_iteration(th, Fo, do)
        P.assert_fp64(F, d, Fo, do, f"chunk over {spans} z bins")
