"""Chunk's protected matrix builders on the device (needs an MI355X).

qml_chunk_matrix and qml_chunk_marginalize against the oracle's restatement
of _setVZMatrices, _setFiducialSignalMatrix, _setQiMatrix, the weights of
_applyRedshiftInterp and _scaleDerivativesWithRedshiftGrowth and
_addMarginalizations (core/chunk_estimate.cpp:384-665), in the table and
resolution-matrix modes; and the friend test of tests/testSQCMatrices.cpp:
24-95, compiled against the drop-in headers (tests/cpp/dropin_check friend),
against the reference's golden signal and Q_0 matrices with its own
allClose bar, the rest of the stored set against the oracle.
"""
import os
import subprocess

import numpy as np
import pytest

import goldencase as gc
import parity as P
from lyspeq_amd import qml, sqio, synth
from lyspeq_amd.build import DROPIN_CHECK
from lyspeq_amd.qml import QMLEngine, QMLError
from oracle.chunk import Chunk as OracleChunk

pytestmark = pytest.mark.gpu

# Bars, as the largest entry error over the largest entry of the oracle's
# matrix.  v, z, the bin weights and growth factors are one or two roundings
# of the same expressions (log, sqrt, pow of the same doubles); S and Q of the
# tables are the same linear interpolation.  With a resolution matrix S and Q
# are R M R^T summed in another order, and Q_k(v) carries the rounding of v
# itself (about c eps = 3e-11 km/s, on either side) through its phase k v, on
# the oversampled grid summed over the taps of R: there the bar is the case's
# rounding floor, the oracle's distance from its own extended-precision run
# (oracle.chunk.Chunk exact=True), as the parity tests take it (parity.py):
# the engine must lie within FLOOR_FACTOR floors of that run and
# FLOOR_FACTOR + 1 of the oracle, or within RMAT_TOL.
EXACT_TOL = 1e-13
TABLE_TOL = 1e-12
RMAT_TOL = 1e-11
MARG_TOL = 1e-11


def _floor_check(worst, what):
    """worst[k] = (engine vs oracle, oracle vs exact, engine vs exact)."""
    for k, (e, f, x) in worst.items():
        assert e < max(RMAT_TOL, (P.FLOOR_FACTOR + 1) * f), (what, k, e, f)
        assert x < max(RMAT_TOL, P.FLOOR_FACTOR * f), (what, k, x, f)


def _case(name):
    print(f"[case] {name}", flush=True)


def _err(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _upper(a):
    return a[np.triu_indices(a.shape[0])]


def _workload(config):
    if config == "golden":
        kw = dict(gc.BINS_KW)
        w = synth.Workload("golden", qml.Binning(**kw), nv=gc.table().nvpoints, v_length=gc.table().v_length,
                           nz_table=gc.table().nzpoints, pd13=gc.PD13)
        t, sp = gc.table(), gc.spectrum()
        w.pairs = [(70000, 20.0)]
        w.deriv = t.deriv[None]
        w.signal = t.signal[None]
        w.chunks.append(dict(wave=sp.wave, delta=sp.flux, ivar=sp.ivar, R=sp.R_fwhm, dv_table=20.0))
        return w
    if config in ("desi", "growth"):
        w = synth.desi_config(nchunks=6)
        if config == "growth":
            w.binning.redshift_growth_on = True
            w.binning.fiducial_redshift_power = 3.6
        return w
    return synth.rmat_config(nchunks=4, n_range=(100, 170), kind=config.split("-")[1])


def _pairs(w, exact=False):
    """(engine, oracle bins, [(chunk id, oracle chunk[, its extended-precision twin])])."""
    eng = QMLEngine(w.binning, precision="fp64")
    accepted = synth.load_engine(eng, w, list(range(len(w.chunks))))
    eng.commit()
    b = P.oracle_bins(w.binning)
    tabs = P.oracle_tables(w)
    out = []
    for i, cid in accepted:
        c = w.chunks[i]
        args = (b, tabs, c["wave"], c["delta"], c["ivar"], c["R"])
        ch = OracleChunk(*args, rmat=P.oracle_rmat(c))
        out.append((cid, ch, OracleChunk(*args, rmat=P.oracle_rmat(c), exact=True)) if exact else (cid, ch))
    return eng, b, out


def _three(got, want, exact):
    return (_err(got, want), _err(want, exact), _err(got, exact))


def _worst(worst, k, t):
    old = worst.get(k)
    worst[k] = t if old is None or t[0] > old[0] else old


@pytest.mark.parametrize("config", ["golden", "desi", "growth", "rmat-dia", "rmat-oversampled"])
def test_builder_matrices_match_oracle(config):
    _case(f"builder matrices {config}")
    w = _workload(config)
    rm = config.startswith("rmat")
    eng, b, pairs = _pairs(w, exact=rm)
    assert len(pairs) >= 1
    worst, floored = {}, {}
    try:
        for p in pairs:
            cid, ch = p[0], p[1]
            v, z = ch.vz_matrices()
            # v and z on the matrix grid: the oversampled one, else the pixels
            gv, gz = eng.chunk_matrix(cid, qml.CMAT_V), eng.chunk_matrix(cid, qml.CMAT_Z)
            if config == "rmat-oversampled":
                fv, fz, _ = ch._fine_vz()
                assert gv.shape == fv.shape and gv.shape[0] > ch.N
                wv, wz = fv, fz
            else:
                assert gv.shape == (ch.N, ch.N)
                wv, wz = v, z
            assert np.array_equal(gv, gv.T) and np.array_equal(gz, gz.T)
            worst["v"] = max(worst.get("v", 0), _err(_upper(gv), _upper(wv)))
            worst["z"] = max(worst.get("z", 0), _err(_upper(gz), _upper(wz)))
            s = eng.chunk_matrix(cid, qml.CMAT_SFID)
            if rm:
                xv, xz = p[2].vz_matrices()
                _worst(floored, "S", _three(s, ch.signal_matrix(v, z), p[2].signal_matrix(xv, xz)))
            else:
                worst["S"] = max(worst.get("S", 0), _err(s, ch.signal_matrix(v, z)))
            for i_kz in ch.stored:
                kn = (i_kz + ch.fisher_index_start) % b.nk
                q = eng.chunk_matrix(cid, qml.CMAT_Q, i_kz)
                if rm:
                    _worst(floored, "Q", _three(q, ch.q_base(v, kn), p[2].q_base(xv, kn)))
                else:
                    worst["Q"] = max(worst.get("Q", 0), _err(q, ch.q_base(v, kn)))
            zm0 = ch.fisher_index_start // b.nk
            for zm in range(zm0, zm0 + len(ch.stored) // ch.kbin_upp):
                worst["w"] = max(worst.get("w", 0), _err(eng.chunk_matrix(cid, qml.CMAT_ZWEIGHT, zm),
                                                         ch.redshift_weights(z, zm)))
                worst["g"] = max(worst.get("g", 0), _err(eng.chunk_matrix(cid, qml.CMAT_GROWTH, zm),
                                                         ch.growth_factors(z, zm)))
    finally:
        eng.close()
    print(f"[err] {config} " + " ".join(f"{k}={v:.2e}" for k, v in worst.items()), flush=True)
    for k, (e, f, x) in floored.items():
        print(f"[err] {config} {k}: vs oracle {e:.2e}, oracle vs extended {f:.2e}, vs extended {x:.2e}", flush=True)
    for k in ("v", "z", "w", "g"):
        assert worst[k] < EXACT_TOL, (k, worst[k])
    if rm:
        _floor_check(floored, config)
    else:
        for k in ("S", "Q"):
            assert worst[k] < TABLE_TOL, (k, worst[k])


@pytest.mark.parametrize("config,orders", [("golden", (1, -1)), ("golden", (1, 1)), ("rmat-dia", (1, -1))])
def test_marginalization_matches_oracle(config, orders):
    """_addMarginalizations on S + I and on a random symmetric positive
    definite matrix: the chunk's orthonormal templates removed by
    Sherman-Morrison updates, against the oracle's SVD + updates."""
    _case(f"marginalization {config} {orders}")
    w = _workload(config)
    w.binning.cont_logl_order, w.binning.cont_lam_order = orders
    eng, b, pairs = _pairs(w)
    rng = np.random.default_rng(11)
    try:
        for cid, ch in pairs:
            v, z = ch.vz_matrices()
            s = ch.signal_matrix(v, z)
            a = rng.standard_normal((ch.N, ch.N)) / np.sqrt(ch.N)
            for cinv in (s + np.eye(ch.N), a @ a.T + 0.5 * np.eye(ch.N)):
                got = eng.chunk_marginalize(cid, cinv)
                want = ch.add_marginalizations(cinv)
                e = _err(got, want)
                print(f"[err] marginalization {config} {orders} {e:.2e}", flush=True)
                assert e < MARG_TOL
                # the templates are projected out: C_marg^-1 u = 0
                u = ch.templates()
                assert np.max(np.abs(got @ u.T)) < 1e-9 * np.max(np.abs(got))
    finally:
        eng.close()


def test_builder_matrix_arguments_are_checked():
    _case("builder matrix argument errors")
    w = synth.desi_config(nchunks=6)
    eng = QMLEngine(w.binning, precision="fp64")
    accepted = synth.load_engine(eng, w, list(range(len(w.chunks))))
    cid = accepted[0][1]
    try:
        with pytest.raises(QMLError):   # before the commit
            eng.chunk_matrix(cid, qml.CMAT_V)
        eng.commit()
        info = eng.chunk_infos[cid]
        nk = w.binning.nk
        with pytest.raises(QMLError):
            eng.chunk_matrix(cid, 9)
        with pytest.raises(QMLError):
            eng.chunk_matrix(cid, qml.CMAT_Q, info.n_q_matrices)
        beyond = [i for i in range(info.n_q_matrices) if (i + info.fisher_index_start) % nk >= info.kbin_upp]
        assert beyond, "the case needs a band above KBIN_UPP"
        with pytest.raises(QMLError):
            eng.chunk_matrix(cid, qml.CMAT_Q, beyond[0])
        with pytest.raises(QMLError):
            eng.chunk_matrix(cid, qml.CMAT_ZWEIGHT, w.binning.nz)
        with pytest.raises(QMLError):
            eng.chunk_matrix(len(accepted) + 5, qml.CMAT_V)
    finally:
        eng.close()


def _run_friend(tmp_path, cfg_extra, files, tdir, tbase, b, mode="friend"):
    cfg = str(tmp_path / "friend.cfg")
    with open(cfg, "w") as f:
        ke, zc = b.kband_edges(), b.zbin_centers()
        f.write(f"kedges {len(ke)} " + " ".join(repr(float(x)) for x in ke) + "\n")
        f.write(f"zcenters {len(zc)} " + " ".join(repr(float(x)) for x in zc) + "\n")
        f.write(f"zwidth {b.z_bin_width!r}\nmethod {b.method}\nlogl_order {b.cont_logl_order}\n")
        f.write(f"lam_order {b.cont_lam_order}\nmax_forest {b.max_forest_length_v!r}\n")
        f.write(f"nchunks {b.number_of_chunks}\nmax_pixel {b.max_pixel_length_v!r}\n")
        f.write(cfg_extra)
        for p in files:
            f.write(f"file {p}\n")
    out = tmp_path / "out"
    out.mkdir()
    r = subprocess.run([DROPIN_CHECK, mode, cfg, str(tdir), tbase, str(out)], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]

    def rd(name):
        a = np.fromfile(str(out / f"{name}.bin"), dtype=np.float64)
        return a
    return rd


def test_friend_test_against_reference_goldens(tmp_path):
    """tests/testSQCMatrices.cpp:24-95 on the reference's test spectrum and
    table (488 pixels, R = 70000, dv = 20 km/s): the signal matrix and the
    weighted Q_0 against tests/truth/signal_matrix.txt and q0_matrix.txt with
    the test's allClose (rtol 1e-5, atol 1e-8), and to the FP64 bar against
    the oracle; v, z, every stored Q after the redshift weights and the
    marginalisation of S + I (one log-lambda template) against the oracle."""
    if not os.path.exists(DROPIN_CHECK):
        pytest.skip("tests/cpp/dropin_check not built (python -m lyspeq_amd.build)")
    _case("friend test, golden chunk")
    bkw = dict(gc.BINS_KW, cont_logl_order=1)
    b = qml.Binning(**bkw)
    rd = _run_friend(tmp_path, "", [os.path.join(gc.GOLDEN, "desilite_seed332298_id0_0_z2.4_lognormal.dat")],
                     gc.GOLDEN, "signal", b)
    N = 488
    s_gold, q0_gold = gc.chunk_matrices()
    sig, q0 = rd("signal").reshape(N, N), rd("q0").reshape(N, N)
    assert gc.all_close(sig, s_gold) and gc.all_close(q0, q0_gold)
    ob = P.oracle_bins(b)
    t, sp = gc.table(), gc.spectrum()
    from oracle.chunk import TableSet
    tabs = [TableSet(70000, 20.0, t.nvpoints, t.v_length, t.nzpoints, t.deriv, t.signal.reshape(-1))]
    ch = OracleChunk(ob, tabs, sp.wave, sp.flux, sp.ivar, sp.R_fwhm)
    assert ch.N == N
    v, z = ch.vz_matrices()
    assert _err(_upper(rd("v").reshape(N, N)), _upper(v)) < EXACT_TOL
    assert _err(_upper(rd("z").reshape(N, N)), _upper(z)) < EXACT_TOL
    assert _err(sig, ch.signal_matrix(v, z)) < TABLE_TOL
    qs = ch.q_matrices(v, z)
    stored = rd("stored").reshape(len(qs), N, N)
    assert _err(q0, qs[0]) < TABLE_TOL
    for k, q in enumerate(qs):
        assert _err(stored[k], q) < TABLE_TOL, k
    cin = rd("cinv_in").reshape(N, N)
    assert _err(rd("cinv_marg").reshape(N, N), ch.add_marginalizations(cin)) < MARG_TOL


def test_chunk_subclass_built_as_the_cross_exposure(tmp_path):
    """A Chunk subclass that builds itself as cross/exposure.hpp:19-70 does
    (the protected Chunk(), _copyQSOFile, _setNQandFisherIndex, its own
    covariance buffer): its _initMatrices registers the chunk with the
    engine, and its signal matrix and the marginalisation of S + I (one
    log-lambda template) match the golden matrix and the oracle."""
    if not os.path.exists(DROPIN_CHECK):
        pytest.skip("tests/cpp/dropin_check not built (python -m lyspeq_amd.build)")
    _case("Chunk subclass built as the cross exposure, golden chunk")
    b = qml.Binning(**dict(gc.BINS_KW, cont_logl_order=1))
    rd = _run_friend(tmp_path, "", [os.path.join(gc.GOLDEN, "desilite_seed332298_id0_0_z2.4_lognormal.dat")],
                     gc.GOLDEN, "signal", b, mode="exposure")
    N = 488
    s_gold, _ = gc.chunk_matrices()
    sig = rd("ex_signal").reshape(N, N)
    assert gc.all_close(sig, s_gold)
    from oracle.chunk import TableSet
    t, sp = gc.table(), gc.spectrum()
    tabs = [TableSet(70000, 20.0, t.nvpoints, t.v_length, t.nzpoints, t.deriv, t.signal.reshape(-1))]
    ch = OracleChunk(P.oracle_bins(b), tabs, sp.wave, sp.flux, sp.ivar, sp.R_fwhm)
    v, z = ch.vz_matrices()
    assert _err(sig, ch.signal_matrix(v, z)) < TABLE_TOL
    cin = rd("ex_cinv_in").reshape(N, N)
    assert _err(rd("ex_cinv_marg").reshape(N, N), ch.add_marginalizations(cin)) < MARG_TOL


def test_friend_test_resolution_matrix_oversampled_with_growth(tmp_path):
    """The friend test's builders on a picca spectrum with an 11-diagonal
    RESOMAT, deconvolved and oversampled x3, with RedshiftGrowthOn: v and z on
    the oversampled grid (its column count), S = R Shat R^T and the stored Q
    after the bin weights and growth factors, against the oracle."""
    if not os.path.exists(DROPIN_CHECK):
        pytest.skip("tests/cpp/dropin_check not built (python -m lyspeq_amd.build)")
    from oracle import reader as oreader
    from oracle.chunk import ResoMatrix, decide_indices
    _case("friend test, oversampled resolution matrix with growth")
    w = synth.rmat_config(nchunks=1, n_range=(150, 220), kind="dia")
    w.binning.redshift_growth_on = True
    w.binning.fiducial_redshift_power = 3.6
    b = P.oracle_bins(w.binning)
    c = w.chunks[0]
    hdr = [("TARGETID", 1000), ("Z", 3.0), ("RA", 0.0), ("DEC", 0.0),
           ("MEANRESO", 299792.458 / c["R"] / 2.35482004503), ("MEANSNR", 5.0)]
    cols = [("LAMBDA", c["wave"], "D"), ("DELTA", c["delta"], "D"), ("IVAR", c["ivar"], "D"),
            ("RESOMAT", np.asarray(c["rmat"]["values"]).T, "D")]
    path = str(tmp_path / "deltas.fits")
    sqio.write_picca_deltas(path, [(hdr, cols)], compress=False)
    tdir = tmp_path / "tables"
    tdir.mkdir()
    (R0, dv0), = w.pairs
    t = sqio.SQTable(nvpoints=w.nv, nzpoints=w.nz_table, v_length=w.v_length, z1=0.0, z_length=0.0,
                     spectrograph_resolution=R0, pixel_width=dv0, k1=0.0, dklin=0.0, dklog=0.0,
                     nklin=w.binning.nk, nklog=0, deriv=w.deriv[0], signal=w.signal[0])
    sqio.write_sq_table(str(tdir / f"sq_R{R0}_dv{dv0:.1f}.fits"), t)
    extra = (f"format picca\nrmat 1\ndeconv 1.0\nosamp 3\nfirst_table {R0} {dv0!r}\n"
             f"growth {w.binning.fiducial_redshift_power!r}\n")
    rd = _run_friend(tmp_path, extra, [f"{path}[1]"], tdir, "sq", w.binning)
    # the oracle's reading of the same spectrum: outliers, the matrix's
    # preprocessing, the boundary cut and the first chunk
    wave = np.asarray(c["wave"], float)
    delta, ivar, _ = oreader.mask_outliers(wave, c["delta"], c["ivar"])
    rows = ResoMatrix("dia", np.asarray(c["rmat"]["values"]), 11).deconvolve(1.0).oversample(3)
    a, e = oreader.cut_boundary(b, wave, delta, ivar)
    ix = decide_indices(wave[a:e], b)
    args = (b, P.oracle_tables(w), wave[a:e], delta[a:e], ivar[a:e], c["R"], ix[0], ix[1])
    ch = OracleChunk(*args, rmat=rows.subset(a, e))
    chx = OracleChunk(*args, rmat=rows.subset(a, e), exact=True)
    N = ch.N
    fv, fz, _ = ch._fine_vz()
    M = fv.shape[0]
    assert M == N * 3 + rows.nelem - 1
    assert _err(_upper(rd("v").reshape(M, M)), _upper(fv)) < EXACT_TOL
    assert _err(_upper(rd("z").reshape(M, M)), _upper(fz)) < EXACT_TOL
    v, z = ch.vz_matrices()
    xv, xz = chx.vz_matrices()
    floored = {}
    _worst(floored, "S", _three(rd("signal").reshape(N, N), ch.signal_matrix(v, z), chx.signal_matrix(xv, xz)))
    qs, xqs = ch.q_matrices(v, z), chx.q_matrices(xv, xz)
    stored = rd("stored").reshape(len(qs), N, N)
    for k, q in enumerate(qs):
        _worst(floored, "Q", _three(stored[k], q, xqs[k]))
    for k, (e, f, x) in floored.items():
        print(f"[err] friend oversampled {k}: vs oracle {e:.2e}, oracle vs extended {f:.2e}, vs extended {x:.2e}",
              flush=True)
    _floor_check(floored, "friend test, oversampled resolution matrix")
    cin = rd("cinv_in").reshape(N, N)
    assert _err(rd("cinv_marg").reshape(N, N), ch.add_marginalizations(cin)) < MARG_TOL
