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
    base = str(tmp_path / "qmle")
    with open(cfg, "a") as f:
        f.write(f"save_chunks {base}\nsave_boot {base}\n")
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
    # _saveChunkResults' file holds every chunk's d, b, t and Fisher upper
    # triangle; the realizations file the bootstrap powers
    saved = sqio.read_boot_chunks(base + "-bootchunks-0.fits")
    assert len(saved) == len(oflat)
    for i, sc in enumerate(saved):
        Fc, dc = gout[i]
        iu = np.triu_indices(Fc.shape[0])
        assert sc.istart == chunks_out[i]["fis"] and sc.d.size == chunks_out[i]["nq"]
        assert np.array_equal(np.array([sc.d, sc.b, sc.t]), dc) and np.array_equal(sc.fisher[iu], Fc[iu])
    bp, binv, bn, bk, bz, bfast = sqio.read_bootstrap_realizations(base + "-bootstrap-realizations.fits")
    assert (bn, bk * bz, bfast) == (nboots, T, fast)
    assert np.array_equal(bp.reshape(-1), rec["bootpower"]) and np.array_equal(binv, F)
    if not fast:
        tp, tf = np.zeros(T), np.zeros((T, T))
        for s_, (lo, n) in enumerate(spans):
            tp[lo:lo + n] += weights[s_][0] * thetas[s_]
            tf[lo:lo + n, lo:lo + n] += weights[s_][0] * fishers[s_]
        np.testing.assert_allclose(rec["boot0power"], tp, rtol=1e-13, atol=0)
        np.testing.assert_allclose(rec["boot0fisher"].reshape(T, T), tf, rtol=1e-13, atol=0)


def _read(tmp_path, fmt, path, name="read.txt"):
    out = str(tmp_path / name)
    r = subprocess.run([DROPIN_CHECK, "read", fmt, path, out], capture_output=True, text=True, timeout=60)
    if r.returncode != 0:
        return None, r.stderr
    rec = {}
    with open(out) as f:
        for ln in f:
            t = ln.split()
            rec[t[0]] = int(t[1]) if len(t) == 2 else np.array(t[2:], float)
    return rec, ""


def test_picca_reader_matches_the_binary_file_of_the_same_spectrum(tmp_path):
    """The reference's own test spectrum in both of its formats
    (tests/input/desilite_seed332298_id0_0_z2.4_lognormal.{fits,dat}): the
    picca reader (LAMBDA, DELTA, IVAR columns; TARGETID, MEANRESO, DLL keys)
    and the binary reader must agree on the pixels, R_fwhm = 70000 and
    dv = 20 km/s (DLL c ln 10 against the header's pixel width)."""
    _need_exe()
    g = os.path.join(os.path.dirname(__file__), "golden", "desilite_seed332298_id0_0_z2.4_lognormal")
    pa, err = _read(tmp_path, "picca", g + ".fits", "a.txt")
    assert pa is not None, err
    bi, err = _read(tmp_path, "binary", g + ".dat", "b.txt")
    assert bi is not None, err
    assert pa["id"] == 1 and pa["R_fwhm"] == bi["R_fwhm"] == 70000
    assert pa["expid"] == pa["night"] == pa["fiber"] == pa["petal"] == -1
    assert pa["wave"].size == 488
    assert np.array_equal(pa["wave"], bi["wave"]) and np.array_equal(pa["delta"], bi["delta"])
    np.testing.assert_allclose(pa["ivar"], bi["ivar"], rtol=1e-14, atol=0)
    # z, ra, dec, R_kms, snr, dv_kms, dlambda
    np.testing.assert_allclose(pa["params"], bi["params"], rtol=1e-14, atol=0)
    assert abs(pa["params"][5] - 20.0) < 1e-12


def _picca_spectrum(rng, n, nd=0, form="lambda", ivar_form="D"):
    wave = 3600.0 * np.exp(np.arange(n) * 20.0 / 299792.458)
    cols = [("LOGLAM", np.log10(wave), "D")] if form == "loglam" else [("LAMBDA", wave, "D")]
    cols += [("DELTA_BLIND" if form == "loglam" else "DELTA", rng.standard_normal(n), "D"),
             ("IVAR", rng.uniform(0.5, 2.0, n), ivar_form)]
    if nd:
        cols.append(("RESOMAT", rng.uniform(0, 1, (n, nd)), "D"))
    return cols


def test_picca_reader_columns_keys_hdus_and_compression(tmp_path):
    """Synthetic picca files (plain and gzip): "file[k]" selects spectrum k,
    a bare name the first; LOGLAM becomes 10^x, DELTA_BLIND stands in for
    DELTA, a float IVAR column is widened, THING_ID for TARGETID, the
    optional keys default to -1, R_fwhm = int(c / MEANRESO / 2.35482004503 /
    100 + 0.5) * 100, dv = DLL c 2.30258509299, and RESOMAT's N rows of
    ndiags values come back as the diagonal storage ndiags x N
    (io/qso_file.cpp:538-677).  Files missing a required key or column are
    refused with the reference's messages."""
    _need_exe()
    rng = np.random.default_rng(5)
    c = 299792.458
    s0 = ([("TARGETID", 39627), ("Z", 2.75), ("RA", 0.5), ("DEC", -0.25), ("MEANRESO", 13.4),
           ("MEANSNR", 3.5), ("DLL", 1e-4)], _picca_spectrum(rng, 150))
    s1 = ([("THING_ID", 91), ("Z", 3.1), ("RA", 1.0), ("DEC", 0.1), ("MEANRESO", 60.0), ("MEANSNR", 1.5),
           ("DLAMBDA", 0.8), ("EXPID", 12), ("NIGHT", 20210105), ("FIBER", 4), ("PETAL_LOC", 7)],
          _picca_spectrum(rng, 211, nd=11, form="loglam", ivar_form="E"))
    s2 = ([("TARGETID", 5), ("Z", 2.2), ("RA", 0.0), ("DEC", 0.0), ("MEANRESO", 20.0), ("MEANSNR", 1.0)],
          [("FLUX", np.ones(120), "D"), ("DELTA", np.zeros(120), "D"), ("IVAR", np.ones(120), "D")])
    s3 = ([("Z", 2.2), ("RA", 0.0), ("DEC", 0.0), ("MEANRESO", 20.0), ("MEANSNR", 1.0)], _picca_spectrum(rng, 120))
    for compress in (False, True):
        path = str(tmp_path / ("deltas.fits.gz" if compress else "deltas.fits"))
        sqio.write_picca_deltas(path, [s0, s1, s2, s3], compress=compress)
        r0, err = _read(tmp_path, "picca", path)
        assert r0 is not None, err
        r0b, _ = _read(tmp_path, "picca", path + "[1]", "r0b.txt")
        assert np.array_equal(r0["wave"], r0b["wave"]) and r0b["id"] == 39627
        cols0 = dict((k, v) for k, v, _ in s0[1])
        assert r0["id"] == 39627 and r0["expid"] == r0["night"] == r0["fiber"] == r0["petal"] == -1
        assert np.array_equal(r0["wave"], cols0["LAMBDA"]) and np.array_equal(r0["delta"], cols0["DELTA"])
        assert np.array_equal(r0["ivar"], cols0["IVAR"]) and "rmat" not in r0
        assert r0["R_fwhm"] == int(c / 13.4 / 2.35482004503 / 100 + 0.5) * 100
        z, ra, dec, rk, snr, dv, dlam = r0["params"]
        assert (z, ra, dec, rk, snr) == (2.75, 0.5, -0.25, 13.4, 3.5)
        assert dv == 1e-4 * c * 2.30258509299
        # dlambda absent: the median pixel step in wavelength
        assert abs(dlam - np.median(np.diff(cols0["LAMBDA"]))) <= 1e-12 * dlam
        r1, err = _read(tmp_path, "picca", path + "[2]", "r1.txt")
        assert r1 is not None, err
        cols1 = dict((k, v) for k, v, _ in s1[1])
        assert r1["id"] == 91 and (r1["expid"], r1["night"], r1["fiber"], r1["petal"]) == (12, 20210105, 4, 7)
        np.testing.assert_allclose(r1["wave"], 10.0 ** cols1["LOGLAM"], rtol=4e-16, atol=0)
        assert np.array_equal(r1["delta"], cols1["DELTA_BLIND"])
        assert np.array_equal(r1["ivar"], cols1["IVAR"].astype(np.float32).astype(float))
        assert r1["params"][6] == 0.8
        # dv absent: the median of c ln(lambda_i+1 / lambda_i)
        w1 = r1["wave"]
        assert abs(r1["params"][5] - np.median(c * np.log(w1[1:] / w1[:-1]))) < 1e-9
        assert r1["rmat_nelem"] == 11
        assert np.array_equal(r1["rmat"].reshape(11, 211), cols1["RESOMAT"].T)
        bad, err = _read(tmp_path, "picca", path + "[3]", "r2.txt")
        assert bad is None and "Wavelength must be in LOGLAM or LAMBDA!" in err
        bad, err = _read(tmp_path, "picca", path + "[4]", "r3.txt")
        assert bad is None and "TARGETID" in err and "THING_ID" in err
        bad, err = _read(tmp_path, "picca", path + "[5]", "r4.txt")
        assert bad is None and "not found" in err


def _set_card(raw: bytes, key: str, value: str, hdu_start: int) -> bytes:
    """raw with the first card `key` at or after hdu_start rewritten to hold value."""
    k = raw.index(key.ljust(8).encode() + b"= ", hdu_start)
    card = (key.ljust(8) + "= " + value.rjust(20)).ljust(80).encode()
    return raw[:k] + card + raw[k + 80:]


def test_picca_reader_refuses_malformed_files(tmp_path):
    """A truncated table, a negative or huge row count, an oversized column
    repeat and a negative axis in an HDU before the wanted one: each file is
    refused with a FITS error, never read past its end."""
    _need_exe()
    rng = np.random.default_rng(9)
    hdr = [("TARGETID", 1), ("Z", 2.5), ("RA", 0.0), ("DEC", 0.0), ("MEANRESO", 13.4), ("MEANSNR", 2.0)]
    path = str(tmp_path / "ok.fits")
    sqio.write_picca_deltas(path, [(hdr, _picca_spectrum(rng, 150, nd=5)), (hdr, _picca_spectrum(rng, 90))],
                            compress=False)
    raw = open(path, "rb").read()
    good, err = _read(tmp_path, "picca", path + "[2]", "good.txt")
    assert good is not None and good["wave"].size == 90, err
    h1 = raw.index(b"XTENSION")            # the first spectrum's header
    h2 = raw.index(b"XTENSION", h1 + 80)   # the second's
    cases = {
        "truncated": raw[:h1 + 2880 + 2880 // 2],
        "negative_rows": _set_card(raw, "NAXIS2", "-5", h1),
        "huge_rows": _set_card(raw, "NAXIS2", "4611686018427387904", h1),
        "huge_repeat": _set_card(raw, "TFORM4", "'99999999999D'", h1),
        "negative_axis_before": _set_card(raw, "NAXIS2", "-7", h1),
    }
    for name, data in cases.items():
        p = str(tmp_path / f"{name}.fits")
        with open(p, "wb") as f:
            f.write(data)
        which = "[2]" if name == "negative_axis_before" else "[1]"
        rec, err = _read(tmp_path, "picca", p + which, f"{name}.txt")
        assert rec is None and "FITS ERROR" in err, (name, err)
    assert h2 > h1


@pytest.mark.gpu
def test_dropin_picca_files_with_resolution_matrices_against_oracle(tmp_path):
    """Spectra in a picca delta file with an 11-diagonal RESOMAT column, read
    with UseResoMatrix, ResoMatrixDeconvolutionM = 1 and OversampleRmat = 3
    (core/one_qso_estimate.cpp:74-89): the drop-in reads, deconvolves and
    oversamples each spectrum's matrix, cuts, chunks, iterates and flushes;
    the oracle does the same to its own reading of the arrays.  FP64 bar,
    with the floor of the reference's own rounding (the oracle against its
    extended-precision run) for these matrices."""
    from oracle.chunk import ResoMatrix
    _need_exe()
    print("[case] drop-in, picca files with RESOMAT, deconvolved and oversampled", flush=True)
    w = synth.rmat_config(nchunks=8, n_range=(150, 320), kind="dia")
    b = P.oracle_bins(w.binning)
    T = b.total_kz_bins
    c_kms = 299792.458
    spectra = []
    for i, c in enumerate(w.chunks):
        hdr = [("TARGETID", 1000 + i), ("Z", 3.0), ("RA", 0.0), ("DEC", 0.0),
               ("MEANRESO", c_kms / c["R"] / 2.35482004503), ("MEANSNR", 5.0)]
        cols = [("LAMBDA", c["wave"], "D"), ("DELTA", c["delta"], "D"), ("IVAR", c["ivar"], "D"),
                ("RESOMAT", np.asarray(c["rmat"]["values"]).T, "D")]
        spectra.append((hdr, cols))
    path = str(tmp_path / "deltas.fits.gz")
    sqio.write_picca_deltas(path, spectra, compress=True)
    files = [f"{path}[{i + 1}]" for i in range(len(spectra))]
    tdir = tmp_path / "tables"
    tdir.mkdir()
    (R0, dv0), = w.pairs
    t = sqio.SQTable(nvpoints=w.nv, nzpoints=w.nz_table, v_length=w.v_length, z1=0.0, z_length=0.0,
                     spectrograph_resolution=R0, pixel_width=dv0, k1=0.0, dklin=0.0, dklog=0.0,
                     nklin=w.binning.nk, nklog=0, deriv=w.deriv[0], signal=w.signal[0])
    sqio.write_sq_table(str(tdir / f"sq_R{R0}_dv{dv0:.1f}.fits"), t)
    tabs = P.oracle_tables(w)
    Fo, do = np.zeros((T, T)), np.zeros((3, T))
    Fx, dx = np.zeros((T, T)), np.zeros((3, T))
    n_oracle = 0
    for c in w.chunks:
        wave = np.asarray(c["wave"], float)
        delta, ivar, _ = oreader.mask_outliers(wave, c["delta"], c["ivar"])
        rows = ResoMatrix("dia", np.asarray(c["rmat"]["values"]), 11).deconvolve(1.0).oversample(3)
        cut = oreader.cut_boundary(b, wave, delta, ivar)
        if cut is None:
            continue
        a, e = cut
        if np.count_nonzero(ivar[a:e]) < oreader.MIN_PIXELS_IN_FILE:
            continue
        rm = rows.subset(a, e)
        ix = decide_indices(wave[a:e], b)
        for k in range(len(ix) - 1):
            args = (b, tabs, wave[a:e], delta[a:e], ivar[a:e], c["R"], ix[k], ix[k + 1])
            try:
                ch = OracleChunk(*args, rmat=rm)
            except ChunkError:
                continue
            assert ch.one_iteration(np.zeros(T), Fo, do)
            OracleChunk(*args, rmat=rm, exact=True).one_iteration(np.zeros(T), Fx, dx)
            n_oracle += 1
    assert n_oracle >= 6
    cfg = str(tmp_path / "run.cfg")
    _write_config(cfg, w.binning, files)
    with open(cfg, "a") as f:
        f.write(f"format picca\nrmat 1\ndeconv 1.0\nosamp 3\nfirst_table {R0} {dv0!r}\n")
    outp = str(tmp_path / "out.txt")
    r = subprocess.run([DROPIN_CHECK, "run", cfg, str(tdir), "sq", "fp64", outp], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    rec, nchunks = {}, 0
    with open(outp) as f:
        for ln in f:
            t = ln.split()
            if t[0] == "chunk":
                nchunks += 1
            elif t[0] not in ("cF", "cdbt", "collapsed"):
                rec[t[0]] = np.array(t[2:], float)
    assert nchunks == n_oracle
    F = rec["fisher"].reshape(T, T)
    dbt = np.array([rec["d"], rec["b"], rec["t"]])
    floor = P.errors(Fo, do, Fx, dx)
    P.assert_fp64(F, dbt, Fo, do, "drop-in, picca files with RESOMAT", floor=floor, exact=(Fx, dx))


def _text_records(path):
    recs = []
    with open(path) as f:
        for ln in f:
            t = ln.split()
            recs.append((t[0], t[1:]))
    return recs


def test_bootstrap_output_files_round_trip(tmp_path):
    """BootstrapChunksFile::writeChunk and saveBootstrapRealizations
    (io/bootstrap_file.cpp:17-185) from fixed arrays: three chunks of 4, 1 and
    6 bands hold d, b, t and the upper diagonals of F with their TARGETID,
    ZQSO, RA, DEC, NQDIM and ISTART; the realizations file holds the powers
    (nboots x nk nz) and the solver matrix with NBOOTS, NK, NZ, FASTBOOT.  The
    C++ reader (readBootstrapRealizations) and this package's readers must
    return the written values bit for bit."""
    _need_exe()
    base = str(tmp_path / "run")
    out = str(tmp_path / "bootio.txt")
    r = subprocess.run([DROPIN_CHECK, "bootio", base, out], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    recs = _text_records(out)
    want, cur = [], None
    for tag, v in recs:
        if tag == "chunk":
            cur = dict(dim=int(v[0]), istart=int(v[1]), id=int(v[2]))
            want.append(cur)
        elif tag in ("par", "p", "n", "t", "F", "powers", "invf"):
            arr = np.array(v[1:], float)
            if cur is not None and tag not in ("powers", "invf"):
                cur[tag] = arr
            else:
                want.append((tag, arr))
    chunks = [w_ for w_ in want if isinstance(w_, dict)]
    got = sqio.read_boot_chunks(base + "-bootchunks-3.fits")
    assert len(got) == len(chunks) == 3
    for g, w_ in zip(got, chunks):
        n = w_["dim"]
        assert g.targetid == w_["id"] and g.istart == w_["istart"] and g.d.size == n
        assert (g.z_qso, g.ra, g.dec) == tuple(w_["par"])
        assert np.array_equal(g.d, w_["p"]) and np.array_equal(g.b, w_["n"]) and np.array_equal(g.t, w_["t"])
        F = w_["F"].reshape(n, n)
        iu = np.triu_indices(n)
        assert np.array_equal(g.fisher[iu], F[iu]) and np.array_equal(g.fisher, g.fisher.T)
    arrays = dict(x for x in want if not isinstance(x, dict))
    p, inv, nb, nk, nz, fast = sqio.read_bootstrap_realizations(base + "-bootstrap-realizations.fits")
    assert (nb, nk, nz, fast) == (5, 3, 2, True)
    assert np.array_equal(p.reshape(-1), arrays["powers"]) and np.array_equal(inv.reshape(-1), arrays["invf"])
