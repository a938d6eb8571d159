"""The drop-in Chunk / OneQSOEstimate (lyspeq_amd/host/dropin), compiled into
tests/cpp/dropin_check against headers that mirror core/chunk_estimate.hpp:
46-119 and core/one_qso_estimate.hpp:16-61 (its static assertions pin every
signature at build time).

CPU: the reference's cost model (Chunk / OneQSOEstimate::getComputeTimeEst)
against the oracle's restatement on DESI-like spectra read from binary
files.  GPU: spectra read from binary files, tables from SQ FITS files, one
iteration through OneQSOEstimate::oneQSOiteration and the flush, against the
oracle; collapseBootstrap and addBootPowerOnly / addBoot against the same
sums restated."""
import os
import subprocess

import numpy as np
import pytest

import parity as P
from lyspeq_amd import sqio, synth
from lyspeq_amd.build import DROPIN_CHECK
from oracle import cost as ocost
from oracle import reader as oreader
from oracle.chunk import Chunk as OracleChunk, ChunkError, decide_indices

def _need_exe():
    if not os.path.exists(DROPIN_CHECK):
        pytest.skip("tests/cpp/dropin_check not built (python -m lyspeq_amd.build)")


def _write_files(tmp_path, w, idx, snr=10.0):
    files = []
    for i in idx:
        c = w.chunks[i]
        iv = np.asarray(c["ivar"], float)
        with np.errstate(divide="ignore"):
            sigma = np.where(iv > 0, 1.0 / np.sqrt(np.where(iv > 0, iv, 1.0)), 1e10)
        spec = sqio.BQSpectrum(z_qso=3.0, dec=0.0, ra=0.0, R_fwhm=int(c["R"]), snr=snr, dv_kms=-1.0,
                               wave=np.asarray(c["wave"], float), flux=np.asarray(c["delta"], float), ivar=iv)
        path = str(tmp_path / f"spec{i}.dat")
        sqio.write_bq_dat(path, spec, sigma)
        files.append(path)
    return files


def _write_config(path, b, files, fast_bootstrap=True, nboots=0, weights=None):
    with open(path, "w") as f:
        ke = b.kband_edges()
        zc = b.zbin_centers()
        f.write(f"kedges {len(ke)} " + " ".join(repr(float(x)) for x in ke) + "\n")
        f.write(f"zcenters {len(zc)} " + " ".join(repr(float(x)) for x in zc) + "\n")
        f.write(f"zwidth {b.z_bin_width!r}\nmethod {b.method}\nlogl_order {b.cont_logl_order}\n")
        f.write(f"lam_order {b.cont_lam_order}\nmax_forest {b.max_forest_length_v!r}\n")
        f.write(f"nchunks {b.number_of_chunks}\nmax_pixel {b.max_pixel_length_v!r}\n")
        f.write(f"fast_bootstrap {1 if fast_bootstrap else 0}\nnboots {nboots}\n")
        if weights is not None:
            f.write(f"weights {len(weights)} " + " ".join(str(int(x)) for x in weights) + "\n")
        for p in files:
            f.write(f"file {p}\n")


def test_cost_model_hand_evaluated():
    """One chunk of the CPU configuration: N = 256 pixels, one top-hat z bin,
    n stored Q = KBIN_UPP.  With x = 2.56 the model is
    n (19 x^3 + x^2 (n + 2)) + 0.45 x^2 (n + 1) + 19 x^3 + n x^2."""
    w = synth.cpu_reference_config(nchunks=1, with_tables=False)
    b = P.oracle_bins(w.binning)
    c = w.chunks[0]
    wave, ivar = np.asarray(c["wave"]), np.asarray(c["ivar"])
    assert wave.size == 256 and ivar[0] != 0 and ivar[-1] != 0
    n = b.max_k_index(np.pi / np.median(np.log(wave[1:] / wave[:-1]) * 299792.458))
    assert n == 19
    # 19 (19 * 16.777216 + 6.5536 * 21) + 0.45 * 6.5536 * 20 + 19 * 16.777216 + 19 * 6.5536
    want = 19 * (318.767104 + 137.6256) + 58.9824 + 318.767104 + 124.5184
    assert abs(ocost.chunk_cost(b, wave, ivar, 0, 256) - want) < 1e-9 * want
    # a short chunk, and a fully masked one, cost nothing
    assert ocost.chunk_cost(b, wave, ivar, 0, 20) == 0.0
    assert ocost.chunk_cost(b, wave, np.zeros_like(ivar), 0, 256) == 0.0


def test_cost_model_cpp_matches_oracle_on_desi_shapes(tmp_path):
    """Chunk::getComputeTimeEst and OneQSOEstimate::getComputeTimeEst of the
    drop-in against the oracle's restatement, on DESI-like spectra (N in
    [150, 700], 13 triangular z bins, Nk = 35), read from binary files."""
    _need_exe()
    w = synth.desi_config(nchunks=60, with_tables=False)
    b = P.oracle_bins(w.binning)
    idx = list(range(len(w.chunks)))
    files = _write_files(tmp_path, w, idx)
    cfg = str(tmp_path / "cost.cfg")
    _write_config(cfg, w.binning, files)
    out = subprocess.run([DROPIN_CHECK, "cost", cfg], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.split("\n")
    per_file, per_chunk = [], []
    for ln in lines:
        t = ln.split()
        if t and t[0] == "file":
            per_file.append((float(t[2]), int(t[3])))
            per_chunk.append([])
        elif t and t[0] == "chunk":
            per_chunk[-1].append((int(t[1]), int(t[2]), float(t[3])))
    assert len(per_file) == len(files)
    nonzero = 0
    for k, path in enumerate(files):
        s = sqio.read_bq_dat(path)
        want, zbin = ocost.spectrum_cost(b, s.wave, s.ivar)
        got, gz = per_file[k]
        assert abs(got - want) <= 1e-12 * max(1.0, abs(want)), (k, got, want)
        if want > 0:
            nonzero += 1
            assert gz == zbin
            cut = oreader.cut_boundary(b, s.wave, s.flux, s.ivar)
            wv, iv = s.wave[cut[0]:cut[1]], s.ivar[cut[0]:cut[1]]
            ix = decide_indices(wv, b)
            assert [(a, e) for a, e, _ in per_chunk[k]] == list(zip(ix[:-1], ix[1:]))
            for a, e, v in per_chunk[k]:
                assert abs(v - ocost.chunk_cost(b, wv, iv, a, e)) <= 1e-12 * max(1.0, v)
    assert nonzero >= 50


@pytest.mark.gpu
@pytest.mark.parametrize("fast", [True, False])
def test_dropin_iteration_and_bootstrap_against_oracle(tmp_path, fast):
    """Spectra in binary files and their (R, dv) tables as SQ FITS files: the
    drop-in reads, chunks, iterates and flushes; the sums and every chunk's
    results must meet the FP64 bar against the oracle on the same reading
    (outliers masked, redshift boundary cut, decideIndices); then
    collapseBootstrap + addBootPowerOnly (and addBoot with the fast form off)
    must equal the same sums of the chunks' own results."""
    _need_exe()
    print(f"[case] drop-in OneQSOEstimate from files, fast bootstrap {fast}", flush=True)
    w = synth.desi_config(nchunks=16)
    b = P.oracle_bins(w.binning)
    T = b.total_kz_bins
    idx = list(range(len(w.chunks)))
    files = _write_files(tmp_path, w, idx)
    tdir = tmp_path / "tables"
    tdir.mkdir()
    for r, (R, dv) in enumerate(w.pairs):
        t = sqio.SQTable(nvpoints=w.nv, nzpoints=w.nz_table, v_length=w.v_length, z1=0.0, z_length=0.0,
                         spectrograph_resolution=R, pixel_width=dv, k1=0.0, dklin=0.0, dklog=0.0,
                         nklin=w.binning.nk, nklog=0, deriv=w.deriv[r], signal=w.signal[r])
        sqio.write_sq_table(str(tdir / f"sq_R{R}_dv{dv:.1f}.fits"), t)
    nboots = 7
    rng = np.random.default_rng(3)
    # the oracle's reading of every file: the spectra that keep a chunk
    tabs = P.oracle_tables(w)
    Fo, do = np.zeros((T, T)), np.zeros((3, T))
    spectra = []
    for path in files:
        s = sqio.read_bq_dat(path)
        r = oreader.read_spectrum(b, s.wave, s.flux, s.ivar, s.snr)
        if r is None:
            continue
        wv, dl, iv = r
        ix = decide_indices(wv, b)
        chunks = []
        for k in range(len(ix) - 1):
            try:
                ch = OracleChunk(b, tabs, wv, dl, iv, s.R_fwhm, ix[k], ix[k + 1])
            except ChunkError:
                continue
            assert ch.one_iteration(np.zeros(T), Fo, do)
            chunks.append(ch)
        if chunks:
            spectra.append(chunks)
    weights = rng.poisson(1.0, (len(spectra), nboots))
    cfg = str(tmp_path / "run.cfg")
    _write_config(cfg, w.binning, files, fast_bootstrap=fast, nboots=nboots, weights=weights.reshape(-1))
    outp = str(tmp_path / "out.txt")
    r = subprocess.run([DROPIN_CHECK, "run", cfg, str(tdir), "sq", "fp64", outp], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    rec, chunks_out, collapsed = {}, [], []
    with open(outp) as f:
        for ln in f:
            t = ln.split()
            if t[0] == "chunk":
                chunks_out.append(dict(s=int(t[1]), fis=int(t[2]), nq=int(t[3]), vecs=[]))
            elif t[0] in ("cF", "cdbt"):
                chunks_out[-1]["vecs"].append(np.array(t[2:], float))
            elif t[0] == "collapsed":
                collapsed.append((int(t[1]), int(t[2]), int(t[3])))
            else:
                rec[t[0]] = np.array(t[2:], float)
    F = rec["fisher"].reshape(T, T)
    dbt = np.array([rec["d"], rec["b"], rec["t"]])
    P.assert_fp64(F, dbt, Fo, do, "drop-in sums")
    oflat = [ch for chunks in spectra for ch in chunks]
    assert len(chunks_out) == len(oflat)
    gout = {i: (c["vecs"][0].reshape(c["nq"], c["nq"]), np.array(c["vecs"][1:])) for i, c in enumerate(chunks_out)}
    oout = {i: (ch.fisher, ch.dbt) for i, ch in enumerate(oflat)}
    for i, ch in enumerate(oflat):
        assert chunks_out[i]["fis"] == ch.fisher_index_start and chunks_out[i]["nq"] == ch.n_q
    P.assert_chunks_fp64(gout, oout, "drop-in chunks")
    # the bootstrapper's sums from the drop-in's own chunk results
    thetas, fishers, spans = [], [], []
    pos = 0
    for chunks in spectra:
        mine = [gout[pos + k] for k in range(len(chunks))]
        lo = min(ch.fisher_index_start for ch in chunks)
        hi = max(ch.fisher_index_start + ch.n_q for ch in chunks)
        th = np.zeros(hi - lo)
        fm = np.zeros((hi - lo, hi - lo))
        for (Fc, dc), ch in zip(mine, chunks):
            o = ch.fisher_index_start - lo
            th[o:o + ch.n_q] += (dc[0] - dc[1]) - dc[2]
            iu = np.triu_indices(ch.n_q)
            blk = np.zeros((ch.n_q, ch.n_q))
            blk[iu] = Fc[iu]
            fm[o:o + ch.n_q, o:o + ch.n_q] += blk
        thetas.append(th)
        fishers.append(fm)
        spans.append((lo, hi - lo))
        pos += len(chunks)
    assert [(a, n) for a, n, _ in collapsed] == spans
    assert all(c == 0 for _, _, c in collapsed)   # the chunks are released
    power = np.zeros((nboots, T))
    for s_, (lo, n) in enumerate(spans):
        power[:, lo:lo + n] += weights[s_][:, None] * thetas[s_][None, :]
    np.testing.assert_allclose(rec["bootpower"].reshape(nboots, T), power, rtol=1e-13, atol=0)
    if not fast:
        tp, tf = np.zeros(T), np.zeros((T, T))
        for s_, (lo, n) in enumerate(spans):
            tp[lo:lo + n] += weights[s_][0] * thetas[s_]
            tf[lo:lo + n, lo:lo + n] += weights[s_][0] * fishers[s_]
        np.testing.assert_allclose(rec["boot0power"], tp, rtol=1e-13, atol=0)
        np.testing.assert_allclose(rec["boot0fisher"].reshape(T, T), tf, rtol=1e-13, atol=0)
