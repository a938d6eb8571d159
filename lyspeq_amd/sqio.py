"""File formats used around the per-chunk QML path.

* SQ lookup tables: a FITS file with three image HDUs, METADATA (header keys
  only), DERIVATIVE and SIGNAL (io/sq_lookup_table_file.cpp:36-171).  The
  DERIVATIVE image is written from an array laid out k-band-major with the
  velocity axis contiguous, ``data[kn * nv + iv]``
  (core/sq_table.cpp:270), although its NAXIS1 is declared as Nk
  (io/sq_lookup_table_file.cpp:121-134).  SIGNAL is laid out with the
  redshift axis contiguous, ``data[iz + nz * iv]`` (core/sq_table.cpp:308).
  Only the raw byte order matters, so both are returned in that order.
* Binary spectrum files (".dat"): an 88-byte header with natural alignment
  (int N; double z, dec, ra; int R; double snr, dv, 4 x lambda) followed by
  lambda[N], flux[N], sigma[N] (io/qso_file.hpp:20-37,
  io/qso_file.cpp:332-369).  sigma is converted to ivar = 1/sigma^2 with
  values below 1e-15 set to 0 (io/qso_file.cpp:362-366).
* Text matrices: "nrows ncols" then the values row by row
  (mathtools/matrix_helper.cpp:462-505).
* picca delta files (written here for tests and for users feeding the
  drop-in reader): one BINTABLE HDU per spectrum after an empty primary.
* Bootstrap outputs (read here): the per-chunk results file and the
  realizations file of io/bootstrap_file.cpp.

The FITS reader handles exactly what the table writer produces (uncompressed
BITPIX=-64 images in 2880-byte blocks), so no CFITSIO is needed.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

_BLOCK = 2880


def _parse_header(buf: bytes, off: int):
    cards = {}
    while True:
        blk = buf[off:off + _BLOCK]
        if len(blk) < _BLOCK:
            raise ValueError("truncated FITS header")
        off += _BLOCK
        done = False
        for i in range(0, _BLOCK, 80):
            card = blk[i:i + 80].decode("ascii")
            key = card[:8].strip()
            if key == "END":
                done = True
                break
            if card[8:10] != "= ":
                continue
            val = card[10:].split("/")[0].strip()
            if val.startswith("'"):
                val = val.strip("'").strip()
            elif val in ("T", "F"):
                val = val == "T"
            else:
                try:
                    val = int(val)
                except ValueError:
                    val = float(val)
            cards[key] = val
        if done:
            return cards, off


def _read_hdus(path: str):
    with open(path, "rb") as f:
        buf = f.read()
    off = 0
    hdus = []
    while off < len(buf):
        cards, off = _parse_header(buf, off)
        naxis = cards.get("NAXIS", 0)
        shape = [cards.get(f"NAXIS{i + 1}", 0) for i in range(naxis)]
        nelem = int(np.prod(shape)) if naxis > 0 else 0
        bitpix = cards.get("BITPIX", -64)
        nbytes = nelem * abs(bitpix) // 8
        data = None
        if nelem > 0:
            if bitpix != -64:
                raise ValueError("only BITPIX=-64 images are supported")
            data = np.frombuffer(buf, dtype=">f8", count=nelem, offset=off).astype(np.float64)
        off += ((nbytes + _BLOCK - 1) // _BLOCK) * _BLOCK
        hdus.append((cards, data))
    return hdus


@dataclass
class SQTable:
    """One (R, dv) lookup table, with the METADATA keys of
    io/sq_lookup_table_file.hpp:22-32."""
    nvpoints: int
    nzpoints: int
    v_length: float
    z1: float
    z_length: float
    spectrograph_resolution: int
    pixel_width: float
    k1: float
    dklin: float
    dklog: float
    nklin: int
    nklog: int
    deriv: np.ndarray   # (nk, nv), velocity contiguous
    signal: np.ndarray  # (nv, nz), redshift contiguous

    @property
    def nk(self) -> int:
        return self.nklin + self.nklog


def read_sq_table(path: str) -> SQTable:
    hdus = _read_hdus(path)
    by_name = {c.get("EXTNAME", ""): (c, d) for c, d in hdus}
    meta = by_name["METADATA"][0]
    nv, nz = int(meta["NVPTS"]), int(meta["NZPTS"])
    nk = int(meta["NKLIN"]) + int(meta["NKLOG"])
    dc, dd = by_name["DERIVATIVE"]
    if dc["NAXIS1"] != nk or dc["NAXIS2"] != nv:
        raise ValueError("DERIVATIVE has inconsistent Nk or Nv")
    sc, sd = by_name["SIGNAL"]
    if sc["NAXIS1"] != nz or sc["NAXIS2"] != nv:
        raise ValueError("SIGNAL has inconsistent Nz or Nv")
    return SQTable(
        nvpoints=nv, nzpoints=nz, v_length=float(meta["LENV"]),
        z1=float(meta["Z1"]), z_length=float(meta["LENZ"]),
        spectrograph_resolution=int(meta["RFWHM"]), pixel_width=float(meta["DV"]),
        k1=float(meta["K1"]), dklin=float(meta["DKLIN"]), dklog=float(meta["DKLOG"]),
        nklin=int(meta["NKLIN"]), nklog=int(meta["NKLOG"]),
        deriv=dd.reshape(nk, nv).copy(), signal=sd.reshape(nv, nz).copy())


def _card(key: str, val) -> bytes:
    if len(key) > 8:   # the HIERARCH convention for longer keywords
        v = repr(val).upper() if isinstance(val, float) else (f"'{val}'" if isinstance(val, str) else int(val))
        return f"HIERARCH {key} = {v}".ljust(80).encode("ascii")
    if isinstance(val, bool):
        v = "T" if val else "F"
        s = f"{key:<8}= {v:>20}"
    elif isinstance(val, int):
        s = f"{key:<8}= {val:>20d}"
    elif isinstance(val, float):
        s = f"{key:<8}= {repr(val).upper():>20}"
    else:
        q = f"'{val:<8}'"
        s = f"{key:<8}= {q:<20}"
    return s.ljust(80).encode("ascii")


def _header(cards) -> bytes:
    out = b"".join(_card(k, v) for k, v in cards) + b"END".ljust(80)
    pad = (-len(out)) % _BLOCK
    return out + b" " * pad


def _image(data: np.ndarray) -> bytes:
    raw = np.ascontiguousarray(data, dtype=">f8").tobytes()
    return raw + b"\0" * ((-len(raw)) % _BLOCK)


def write_sq_table(path: str, t: SQTable) -> None:
    """Writes the three-HDU layout read by io/sq_lookup_table_file.cpp."""
    meta = [("SIMPLE", True), ("BITPIX", -64), ("NAXIS", 1), ("NAXIS1", 0),
            ("EXTEND", True), ("EXTNAME", "METADATA"),
            ("NVPTS", t.nvpoints), ("NZPTS", t.nzpoints), ("LENV", float(t.v_length)),
            ("Z1", float(t.z1)), ("LENZ", float(t.z_length)),
            ("RFWHM", int(t.spectrograph_resolution)), ("DV", float(t.pixel_width)),
            ("K1", float(t.k1)), ("DKLIN", float(t.dklin)), ("DKLOG", float(t.dklog)),
            ("NKLIN", int(t.nklin)), ("NKLOG", int(t.nklog))]
    deriv = [("XTENSION", "IMAGE"), ("BITPIX", -64), ("NAXIS", 2), ("NAXIS1", t.nk),
             ("NAXIS2", t.nvpoints), ("PCOUNT", 0), ("GCOUNT", 1), ("EXTNAME", "DERIVATIVE")]
    sig = [("XTENSION", "IMAGE"), ("BITPIX", -64), ("NAXIS", 2), ("NAXIS1", t.nzpoints),
           ("NAXIS2", t.nvpoints), ("PCOUNT", 0), ("GCOUNT", 1), ("EXTNAME", "SIGNAL")]
    with open(path, "wb") as f:
        f.write(_header(meta))
        f.write(_header(deriv))
        f.write(_image(t.deriv))
        f.write(_header(sig))
        f.write(_image(t.signal))


_BINTABLE_FORMS = {"D": ">f8", "E": ">f4", "K": ">i8", "J": ">i4"}


def _bintable(header, columns) -> bytes:
    """One BINTABLE HDU: header (key, value) pairs after the structural keys;
    columns (name, array, form) with form in D, E, K, J and arrays of shape
    (N,) or (N, m) (a vector column of m values per row)."""
    n = len(columns[0][1])
    cols = []
    for name, arr, form in columns:
        a = np.asarray(arr, dtype=_BINTABLE_FORMS[form])
        if a.shape[0] != n:
            raise ValueError(f"column {name}: {a.shape[0]} rows, not {n}")
        cols.append((name, a.reshape(n, -1), form))
    row = np.zeros(n, dtype=[(f"c{i}", c[1].dtype, (c[1].shape[1],)) for i, c in enumerate(cols)])
    for i, c in enumerate(cols):
        row[f"c{i}"] = c[1]
    cards = [("XTENSION", "BINTABLE"), ("BITPIX", 8), ("NAXIS", 2), ("NAXIS1", row.dtype.itemsize),
             ("NAXIS2", n), ("PCOUNT", 0), ("GCOUNT", 1), ("TFIELDS", len(cols))]
    for i, (name, a, form) in enumerate(cols):
        m = a.shape[1]
        cards += [(f"TTYPE{i + 1}", name), (f"TFORM{i + 1}", f"{m}{form}" if m > 1 else form)]
        if m > 1:
            cards.append((f"TDIM{i + 1}", f"({m})"))
    raw = row.tobytes()
    return _header(cards + list(header)) + raw + b"\0" * ((-len(raw)) % _BLOCK)


def write_picca_deltas(path: str, spectra, compress: bool = False) -> None:
    """A picca delta file as io/qso_file.cpp:376-677 reads it: an empty
    primary HDU, then one BINTABLE per spectrum.  Each spectrum is a
    (header, columns) pair for ``_bintable``: header keys such as TARGETID or
    THING_ID, Z, RA, DEC, MEANRESO, MEANSNR, DLL, DLAMBDA; columns LOGLAM or
    LAMBDA, DELTA or DELTA_BLIND, IVAR and optionally RESOMAT (N x ndiags).
    "path[k]" then names spectrum k (HDU k + 1).  compress=True gzips it."""
    import gzip
    data = _header([("SIMPLE", True), ("BITPIX", 8), ("NAXIS", 0), ("EXTEND", True)])
    data += b"".join(_bintable(h, c) for h, c in spectra)
    with (gzip.open(path, "wb") if compress else open(path, "wb")) as f:
        f.write(data)


@dataclass
class ChunkResult:
    """One HDU of a <base>-bootchunks-<pe>.fits file (written by
    _saveChunkResults, core/quadratic_estimate.cpp:33-55, through
    BootstrapChunksFile::writeChunk, io/bootstrap_file.cpp:134-185)."""
    targetid: int
    z_qso: float
    ra: float
    dec: float
    istart: int
    d: np.ndarray
    b: np.ndarray
    t: np.ndarray
    fisher: np.ndarray   # ndim x ndim, symmetric, from the stored upper diagonals


def read_boot_chunks(path: str):
    """Every chunk of a bootchunks file: [d, b, t, diagonal 0, ..., diagonal
    ndim - 1 of the upper triangle] per HDU; an empty primary means none."""
    out = []
    for cards, data in _read_hdus(path):
        if data is None or "NQDIM" not in cards:
            continue
        n = int(cards["NQDIM"])
        if data.size != 3 * n + n * (n + 1) // 2:
            raise ValueError(f"chunk HDU of {data.size} values for NQDIM = {n}")
        F = np.zeros((n, n))
        pos = 3 * n
        for d in range(n):
            idx = np.arange(n - d)
            F[idx, idx + d] = data[pos:pos + n - d]
            F[idx + d, idx] = data[pos:pos + n - d]
            pos += n - d
        out.append(ChunkResult(int(cards["TARGETID"]), float(cards["ZQSO"]), float(cards["RA"]),
                               float(cards["DEC"]), int(cards["ISTART"]), data[:n].copy(),
                               data[n:2 * n].copy(), data[2 * n:3 * n].copy(), F))
    return out


def read_bootstrap_realizations(path: str):
    """(powers (nboots, nk * nz), solver matrix (nk * nz, nk * nz), nboots,
    nk, nz, fast) of a <base>-bootstrap-realizations.fits file
    (io/bootstrap_file.cpp:17-120)."""
    hdus = {c.get("EXTNAME"): (c, d) for c, d in _read_hdus(path)}
    c, p = hdus["REALIZATIONS"]
    _, inv = hdus["SOLV_INVF"]
    nb, nk, nz = int(c["NBOOTS"]), int(c["NK"]), int(c["NZ"])
    return (p.reshape(nb, nk * nz), inv.reshape(nk * nz, nk * nz), nb, nk, nz, bool(c["FASTBOOT"]))


def sq_table_filename(directory: str, base: str, r: int, dv: float) -> str:
    """io/sq_lookup_table_file.cpp:12-21."""
    return f"{directory}/{base}_R{int(r)}_dv{dv:.1f}.fits"


_BQ_HEADER = struct.Struct("<i4xdddi4xdddddd")  # 88 bytes


@dataclass
class BQSpectrum:
    z_qso: float
    dec: float
    ra: float
    R_fwhm: int
    snr: float
    dv_kms: float
    wave: np.ndarray
    flux: np.ndarray
    ivar: np.ndarray


def read_bq_dat(path: str) -> BQSpectrum:
    with open(path, "rb") as f:
        buf = f.read()
    n, z, dec, ra, r, snr, dv, _, _, _, _ = _BQ_HEADER.unpack_from(buf, 0)
    arr = np.frombuffer(buf, dtype="<f8", count=3 * n, offset=_BQ_HEADER.size)
    wave, flux, sig = arr[:n].copy(), arr[n:2 * n].copy(), arr[2 * n:].copy()
    with np.errstate(divide="ignore"):
        ivar = 1.0 / (sig * sig)
    ivar[ivar < 1e-15] = 0
    return BQSpectrum(z, dec, ra, int(r), snr, dv, wave, flux, ivar)


def write_bq_dat(path: str, spec: BQSpectrum, sigma: np.ndarray) -> None:
    n = spec.wave.size
    hdr = _BQ_HEADER.pack(n, spec.z_qso, spec.dec, spec.ra, int(spec.R_fwhm), spec.snr,
                          spec.dv_kms, spec.wave[0], spec.wave[-1], spec.wave[0], spec.wave[-1])
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(np.ascontiguousarray(spec.wave, "<f8").tobytes())
        f.write(np.ascontiguousarray(spec.flux, "<f8").tobytes())
        f.write(np.ascontiguousarray(sigma, "<f8").tobytes())


def read_text_matrix(path: str) -> np.ndarray:
    with open(path) as f:
        nr, nc = (int(x) for x in f.readline().split())
        vals = np.array(f.read().split(), dtype=np.float64)
    if vals.size != nr * nc:
        raise ValueError(f"{path}: expected {nr * nc} values, found {vals.size}")
    return vals.reshape(nr, nc)


def read_rdv_list(path: str):
    """io/io_helper_functions.cpp:136-156: a count, then "R dv" pairs."""
    with open(path) as f:
        toks = f.read().split()
    n = int(toks[0])
    return [(int(toks[1 + 2 * i]), float(toks[2 + 2 * i])) for i in range(n)]
