"""The C++ host layer (lyspeq_amd/host/qml_host.hpp: Chunk, OneQSOEstimate,
Engine::flush) driven the way the reference's driver drives Chunk and
OneQSOEstimate (core/quadratic_estimate.cpp:437-457), through the compiled
test driver tests/cpp/host_driver.  Needs an MI355X.

The C++ path must reproduce the Python binding's results bit for bit (same
library, same chunks) and the oracle to the FP64 bar, including spectra
that OneQSOEstimate::decideIndices splits into several chunks."""
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

import parity as P
from lyspeq_amd import synth
from lyspeq_amd.build import HOST_DRIVER
from lyspeq_amd.qml import PRECISIONS
from oracle.chunk import Chunk as OracleChunk, ChunkError, decide_indices

pytestmark = pytest.mark.gpu



def _write_workload(path, w, theta, precision, iterations=1, prep=None):
    prep = prep or {}
    b = w.binning
    edges, zc = b.kband_edges(), b.zbin_centers()
    nk, nz = len(edges) - 1, len(zc)
    with open(path, "wb") as f:
        f.write(struct.pack("<iii", 0x574C4D51, nk, nz))
        f.write(np.asarray(edges, "<f8").tobytes())
        f.write(np.asarray(zc, "<f8").tobytes())
        f.write(struct.pack("<d", b.z_bin_width))
        f.write(struct.pack("<iiii", b.method, b.cont_logl_order, b.cont_lam_order, int(bool(b.turn_off_sfid))))
        f.write(struct.pack("<did", b.max_forest_length_v, b.number_of_chunks, b.max_pixel_length_v))
        f.write(struct.pack("<id", int(prep.get("oversampling", -1)), float(prep.get("deconvolution_m", -1.0))))
        f.write(struct.pack("<ii", PRECISIONS[precision], iterations))
        f.write(struct.pack("<i", len(w.pairs)))
        for i, (R, dv) in enumerate(w.pairs):
            deriv = np.ascontiguousarray(w.deriv[i], "<f8")
            sig = np.ascontiguousarray(w.signal[i], "<f8")
            f.write(struct.pack("<idid", int(R), float(dv), deriv.shape[1], float(w.v_length)))
            f.write(struct.pack("<i", sig.shape[1]))
            f.write(deriv.tobytes())
            f.write(sig.tobytes())
        f.write(struct.pack("<i", len(w.chunks)))
        for c in w.chunks:
            n = len(c["wave"])
            f.write(struct.pack("<ii", int(c["R"]), n))
            for key in ("wave", "delta", "ivar"):
                f.write(np.ascontiguousarray(c[key], "<f8").tobytes())
            if "rmat" in c:
                assert c["rmat"]["kind"] == "dia"
                f.write(struct.pack("<i", int(c["rmat"]["nelem"])))
                f.write(np.ascontiguousarray(c["rmat"]["values"], "<f8").tobytes())
            else:
                f.write(struct.pack("<i", 0))
        f.write(np.ascontiguousarray(theta, "<f8").tobytes())


def _read_output(path):
    with open(path, "rb") as f:
        buf = f.read()
    T = struct.unpack_from("<i", buf, 0)[0]
    o = 4
    F = np.frombuffer(buf, "<f8", T * T, o).reshape(T, T)
    o += 8 * T * T
    dbt = np.frombuffer(buf, "<f8", 3 * T, o).reshape(3, T)
    o += 8 * 3 * T
    nch = struct.unpack_from("<i", buf, o)[0]
    o += 4
    chunks = []
    for _ in range(nch):
        s, fis, nq, st = struct.unpack_from("<iiii", buf, o)
        o += 16
        fc = np.frombuffer(buf, "<f8", nq * nq, o).reshape(nq, nq)
        o += 8 * nq * nq
        d = np.frombuffer(buf, "<f8", 3 * nq, o).reshape(3, nq)
        o += 8 * 3 * nq
        chunks.append((s, fis, nq, st, fc, d))
    return F, dbt, chunks


def _run_driver(tmp_path, w, theta, precision, iterations=1, prep=None):
    if not os.path.exists(HOST_DRIVER):
        pytest.fail(f"{HOST_DRIVER} is missing: build it with `python -m lyspeq_amd.build`")
    wl, out = tmp_path / "workload.bin", tmp_path / "out.bin"
    _write_workload(wl, w, theta, precision, iterations, prep)
    r = subprocess.run([HOST_DRIVER, str(wl), str(out)], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    return _read_output(out)


def _pieces(w):
    """Chunk pieces (spectrum index, i1, i2) by the oracle's decideIndices."""
    b = P.oracle_bins(w.binning)
    out = []
    for s, c in enumerate(w.chunks):
        idx = decide_indices(np.asarray(c["wave"]), b)
        out += [(s, idx[k], idx[k + 1]) for k in range(len(idx) - 1)]
    return out


def _python_engine(w, precision):
    """The same chunks registered through the Python binding and committed."""
    from lyspeq_amd.qml import QMLEngine, QMLError
    eng = QMLEngine(w.binning, precision=precision)
    for i, (R, dv) in enumerate(w.pairs):
        eng.add_table(R, dv, w.deriv[i], w.signal[i], w.v_length)
    ids = []
    for s, i1, i2 in _pieces(w):
        c = w.chunks[s]
        try:
            cid, _ = eng.add_chunk(c["wave"][i1:i2], c["delta"][i1:i2], c["ivar"][i1:i2], c["R"])
            ids.append(cid)
        except QMLError as e:
            if e.code not in (-3, -4, -5):
                raise
    eng.commit()
    return eng, ids


def _python_path(w, theta, precision):
    """One iteration of the same chunks through the Python binding."""
    eng, ids = _python_engine(w, precision)
    T = eng.T
    F, d = np.zeros((T, T)), np.zeros((3, T))
    eng.iterate(theta, F, d)
    res = [eng.chunk_result(cid) for cid in ids]
    eng.close()
    return F, d, res


def _oracle(w, theta):
    b = P.oracle_bins(w.binning)
    tabs = P.oracle_tables(w)
    T = b.total_kz_bins
    F, d = np.zeros((T, T)), np.zeros((3, T))
    n = 0
    for s, i1, i2 in _pieces(w):
        c = w.chunks[s]
        try:
            ch = OracleChunk(b, tabs, c["wave"][i1:i2], c["delta"][i1:i2], c["ivar"][i1:i2], c["R"])
        except ChunkError:
            continue
        ch.one_iteration(theta, F, d)
        n += 1
    return F, d, n


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_cpp_host_layer_matches_python_binding_and_oracle(tmp_path, precision):
    print(f"[case] C++ host layer, desi sample, {precision}", flush=True)
    w = synth.desi_config(nchunks=16)
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 3)
    F, dbt, chunks = _run_driver(tmp_path, w, th, precision, iterations=2)
    Fp, dp, res = _python_path(w, th, precision)
    assert len(chunks) == len(res) >= 12
    assert np.array_equal(F, Fp) and np.array_equal(dbt, dp)
    for (s, fis, nq, st, fc, d), (fr, dr, sr) in zip(chunks, res):
        assert st == sr == 0
        assert np.array_equal(fc, fr) and np.array_equal(d, dr)
    if precision == "fp64":
        Fo, do, n = _oracle(w, th)
        assert n == len(chunks)
        P.assert_fp64(F, dbt, Fo, do, "C++ host layer")


def test_cpp_one_qso_estimate_splits_spectra_like_decide_indices(tmp_path):
    print("[case] C++ OneQSOEstimate with DynamicChunkNumber 2", flush=True)
    w = synth.cpu_reference_config(nchunks=6)
    # 256 pixels x 60 km/s = 15300 km/s per spectrum; two chunks of ~128 pixels
    w.binning.number_of_chunks = 2
    w.binning.max_forest_length_v = 16000.0
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 9)
    F, dbt, chunks = _run_driver(tmp_path, w, th, "fp64")
    pieces = _pieces(w)
    assert len(pieces) == 2 * len(w.chunks)
    Fp, dp, res = _python_path(w, th, "fp64")
    assert len(chunks) == len(res) == len(pieces)
    assert [c[0] for c in chunks] == [p[0] for p in pieces]
    assert np.array_equal(F, Fp) and np.array_equal(dbt, dp)
    Fo, do, n = _oracle(w, th)
    assert n == len(chunks)
    P.assert_fp64(F, dbt, Fo, do, "C++ OneQSOEstimate split spectra")


def test_cpp_resolution_matrix_spectra_preprocessed_like_the_reader(tmp_path):
    """Spectra carrying an 11-diagonal matrix in diagonal storage, with
    ResoMatrixDeconvolutionM = 1 and OversampleRmat = 3: the C++ host layer
    preprocesses each spectrum (OneQSOEstimate::prepareResolution) and
    registers the chunks' row slices; the Python binding does the same through
    qml_rmat_deconvolve/qml_rmat_oversample; the oracle preprocesses its own
    copy (core/one_qso_estimate.cpp:74-89)."""
    from lyspeq_amd.qml import QMLEngine, QMLError, rmat_deconvolve, rmat_oversample
    from oracle.chunk import ResoMatrix
    print("[case] C++ host layer, resolution matrices deconvolved and oversampled", flush=True)
    w = synth.rmat_config(nchunks=8, n_range=(110, 180), kind="dia")
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 41)
    F, dbt, chunks = _run_driver(tmp_path, w, th, "fp64", prep={"oversampling": 3, "deconvolution_m": 1.0})
    # the same through the Python binding
    eng = QMLEngine(w.binning, precision="fp64")
    for i, (R, dv) in enumerate(w.pairs):
        eng.add_table(R, dv, w.deriv[i], w.signal[i], w.v_length)
    b = P.oracle_bins(w.binning)
    tabs = P.oracle_tables(w)
    T = b.total_kz_bins
    Fo, do = np.zeros((T, T)), np.zeros((3, T))
    Fx, dx = np.zeros((T, T)), np.zeros((3, T))   # the oracle's extended-precision run
    ids, n_oracle = [], 0
    for s, c in enumerate(w.chunks):
        dia = np.asarray(c["rmat"]["values"])
        rows = rmat_oversample(rmat_deconvolve(dia, 1.0), 3)
        orows = ResoMatrix("dia", dia, 11).deconvolve(1.0).oversample(3).values
        idx = decide_indices(np.asarray(c["wave"]), b)
        for k in range(len(idx) - 1):
            i1, i2 = idx[k], idx[k + 1]
            try:
                cid, _ = eng.add_chunk_rmat(c["wave"][i1:i2], c["delta"][i1:i2], c["ivar"][i1:i2], c["R"],
                                            "oversampled", rows[i1:i2], 31, 3)
                ids.append(cid)
            except QMLError as e:
                if e.code not in (-3, -4, -5):
                    raise
            try:
                ch = OracleChunk(b, tabs, c["wave"][i1:i2], c["delta"][i1:i2], c["ivar"][i1:i2], c["R"],
                                 rmat=ResoMatrix("oversampled", orows[i1:i2], 31, 3))
            except ChunkError:
                continue
            ch.one_iteration(th, Fo, do)
            OracleChunk(b, tabs, c["wave"][i1:i2], c["delta"][i1:i2], c["ivar"][i1:i2], c["R"],
                        rmat=ResoMatrix("oversampled", orows[i1:i2], 31, 3), exact=True).one_iteration(th, Fx, dx)
            n_oracle += 1
    eng.commit()
    Fp, dp = np.zeros((T, T)), np.zeros((3, T))
    eng.iterate(th, Fp, dp)
    res = [eng.chunk_result(cid) for cid in ids]
    eng.close()
    assert len(chunks) == len(res) == n_oracle >= 6
    assert np.array_equal(F, Fp) and np.array_equal(dbt, dp)
    for (s, fis, nq, st, fc, d), (fr, dr, sr) in zip(chunks, res):
        assert st == sr == 0 and np.array_equal(fc, fr) and np.array_equal(d, dr)
    floor = P.errors(Fo, do, Fx, dx)
    P.assert_fp64(F, dbt, Fo, do, "preprocessed matrices, C++ layer", floor=floor, exact=(Fx, dx))


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_cpp_newton_raphson_loop_with_graph_replay_and_autotuned_schedule(tmp_path, precision):
    """PowerEstimate (lyspeq_amd/host/qml_driver.hpp) runs three Newton-Raphson
    iterations, each theta feeding the next, over an Engine with graph replay
    and an autotuned batch schedule.  Iteration k is repeated through the
    Python binding with the C++ theta of iteration k - 1 as its fits, so both
    sides hold bit-identical sums.  The C++ theta must then solve the
    reference's equations, F_s theta = (d - b - t) / 2, or with damping
    (F_s F_s + damp I) theta = F_s (d - b - t) / 2, with a backward error no
    worse than 100 times that of the Python driver's theta (LAPACK LU) or
    1e-13, whichever is larger: a bar set by the arithmetic, not by the
    conditioning of F_s.  Both must agree within 1e-3 sigma.  The workload is
    the 64-chunk CPU configuration (20 bands): a few chunks spread over the
    455 bands of the DESI binning leave F_s near-singular, and the first
    iteration's theta then makes no covariance positive definite."""
    from lyspeq_amd.driver import power_estimates
    print(f"[case] C++ Newton-Raphson loop, graph + autotune, {precision}", flush=True)
    w = synth.cpu_reference_config(nchunks=64)
    T = w.binning.total_kz_bins
    iters = 3
    wl, out = tmp_path / "workload.bin", tmp_path / "out.bin"
    _write_workload(wl, w, np.zeros(T), precision, iters)
    r = subprocess.run([HOST_DRIVER, str(wl), str(out), "graph", "autotune", "nr"], capture_output=True,
                       text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    buf = out.read_bytes()
    T2, run = struct.unpack_from("<ii", buf, 0)
    assert T2 == T and run == iters
    hist = np.frombuffer(buf, "<f8", run * T, 8).reshape(run, T)
    sigma = np.frombuffer(buf, "<f8", T, 8 + 8 * run * T)
    eng, ids = _python_engine(w, precision)
    for it in range(iters):
        fits = np.zeros(T) if it == 0 else hist[it - 1]
        F, d = np.zeros((T, T)), np.zeros((3, T))
        eng.iterate(fits, F, d)
        assert sum(eng.chunk_result(cid)[2] == 0 for cid in ids) >= 60
        res = power_estimates(F, d)
        A = res["fisher"].copy()
        A[np.diag(A) == 0, np.diag(A) == 0] = 1.0
        rhs = 0.5 * (d[0] - d[1] - d[2])
        th = hist[it]
        if res["damp"] != 0:
            lhs, want = (A @ A + res["damp"] * np.eye(T)) @ th, A @ rhs
            scale = np.abs(A @ A).max() * np.abs(th).max() + np.abs(want).max()
        else:
            lhs, want = A @ th, rhs
            scale = np.abs(A).max() * np.abs(th).max() + np.abs(want).max()
        backward = np.abs(lhs - want).max() / scale
        thp = res["theta"]
        lhs_p = (A @ A + res["damp"] * np.eye(T)) @ thp if res["damp"] != 0 else A @ thp
        backward_p = np.abs(lhs_p - want).max() / scale
        ok = np.isfinite(res["sigma"]) & (res["sigma"] > 0)
        dev = np.abs(th - thp)[ok] / res["sigma"][ok]
        print(f"iteration {it}: damp {res['damp']:.2e}, backward error C++ {backward:.2e} "
              f"Python {backward_p:.2e}, max |d theta| / sigma {dev.max():.2e}", flush=True)
        assert backward < max(100 * backward_p, 1e-13)
        assert dev.max() < 1e-3
    assert np.allclose(sigma[ok], res["sigma"][ok], rtol=1e-4, atol=0)
    assert not np.allclose(hist[0], hist[1])   # theta moved between iterations
    eng.close()


def test_cpp_bootstrap_over_the_queue_matches_python_binding(tmp_path):
    """Engine::bootstrap numbers the spectra that hold chunks in queue order
    and maps each chunk to its spectrum; the same mapping through the Python
    binding must give the same mean and covariance bit for bit, and
    Engine::bootstrapMedian of those realisations the oracle's median form."""
    from oracle import bootstrap as ob
    from lyspeq_amd.qml import QMLEngine, QMLError
    print("[case] C++ bootstrap over the queue", flush=True)
    w = synth.desi_config(nchunks=16)
    T = w.binning.total_kz_bins
    th = P.theta_perturbed(w.binning, w.pd13, 0.05, 13)
    wl, out = tmp_path / "workload.bin", tmp_path / "out.bin"
    _write_workload(wl, w, th, "fp64", 1)
    r = subprocess.run([HOST_DRIVER, str(wl), str(out), "boot"], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    buf = np.fromfile(str(out) + ".boot", "<f8")
    mean_c, cov_c = buf[:T], buf[T:T + T * T].reshape(T, T)
    med_c, mad_c = buf[T + T * T:2 * T + T * T], buf[2 * T + T * T:].reshape(T, T)
    eng = QMLEngine(w.binning, precision="fp64")
    for i, (R, dv) in enumerate(w.pairs):
        eng.add_table(R, dv, w.deriv[i], w.signal[i], w.v_length)
    spec_of = []
    for s, i1, i2 in _pieces(w):
        c = w.chunks[s]
        try:
            eng.add_chunk(c["wave"][i1:i2], c["delta"][i1:i2], c["ivar"][i1:i2], c["R"])
            spec_of.append(s)
        except QMLError as e:
            if e.code not in (-3, -4, -5):
                raise
    eng.commit()
    F, d = np.zeros((T, T)), np.zeros((3, T))
    eng.iterate(th, F, d)
    order = {s: k for k, s in enumerate(sorted(set(spec_of)))}
    qso = np.array([order[s] for s in spec_of], dtype=np.int32)
    R, mean, cov = eng.bootstrap(500, qso, len(order), seed=7, realizations=True)
    eng.close()
    assert np.array_equal(mean, mean_c) and np.array_equal(cov, cov_c)
    omed, omad = ob.median_covariance(R)
    assert np.array_equal(med_c, omed) and np.array_equal(mad_c, omad)
    assert np.abs(cov).max() > 0
