"""Synthetic chunk batches and S/Q lookup tables for the benchmark configurations.

Tables follow the reference generator's definitions
(core/sq_table.cpp:201-323, core/fiducial_cosmology.cpp:273-331):
  Q_kn(v)  = (1/pi) int_{k_n}^{k_n+1} W(k)^2 cos(k v) dk,
  W(k)     = sinc(k dv/2) exp(-(k R)^2/2), R = c / R_fwhm / 2.35482,
  S(z, v)  = c2r FFT of W^2 P_fid(k, z) on the reference's grid.
The derivative table is integrated with Filon's rule (Simpson-type, exact
for the cosine) on 2M panels per k band, written as one matrix product per
band so thousands of (R, dv) tables cost a few GEMMs.  The signal table uses
the reference's FFT grid exactly.

Chunks: flux contrast is a Gaussian realisation of P_fid smoothed by the
spectrograph Gaussian and the pixel top hat, drawn by FFT on a fine velocity
grid and sampled at the pixel positions; noise sigma is log-normal; a fraction
of interior pixels is masked with ivar = 0.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from .qml import Binning

SPEED_OF_LIGHT = 299792.458
LYA_REST = 1215.67
ONE_SIGMA_2_FWHM = 2.35482004503
MY_PI = 3.14159265358979323846

PD13_DEFAULT = (0.066, -2.7, -0.1, 3.6, 0.0, 1.0 / 0.07 ** 2)


def pd13_power(k, z, A, n, alpha, B, beta, lam):
    """Palanque_Delabrouille_etal_2013_fit (core/fiducial_cosmology.cpp:273-282)."""
    q0 = np.asarray(k) / 0.009 + 1e-15
    x0 = (1.0 + z) / 4.0
    lq = np.log(q0)
    return (A * MY_PI / 0.009 * np.power(q0, 2.0 + n + alpha * lq + beta * np.log(x0))
            * np.power(x0, B) / (1.0 + lam * np.asarray(k) ** 2))


def window_sq(k, dv_kms, R_kms):
    k = np.asarray(k, dtype=np.float64)
    x = k * np.asarray(dv_kms) / 2.0
    s = np.where(x == 0, 1.0, np.sin(x) / np.where(x == 0, 1.0, x))
    kr = k * np.asarray(R_kms)
    w = s * np.exp(-kr * kr / 2.0)
    return w * w


def _filon_coeffs(theta):
    th = np.asarray(theta, dtype=np.float64)
    small = np.abs(th) < 0.02
    t = np.where(small, 1.0, th)
    s, c = np.sin(t), np.cos(t)
    a = (t * t + t * s * c - 2 * s * s) / t ** 3
    b = 2 * (t * (1 + c * c) - 2 * s * c) / t ** 3
    g = 4 * (s - t * c) / t ** 3
    t2 = th * th
    a_s = 2 * th ** 3 / 45 - 2 * th ** 5 / 315 + 2 * th ** 7 / 4725
    b_s = 2 / 3 + 2 * t2 / 15 - 4 * t2 ** 2 / 105 + 2 * t2 ** 3 / 567
    g_s = 4 / 3 - 2 * t2 / 15 + t2 ** 2 / 210 - t2 ** 3 / 11340
    return np.where(small, a_s, a), np.where(small, b_s, b), np.where(small, g_s, g)


def filon_cos_weights(a: float, b: float, t: np.ndarray, M: int):
    """Weights w (len(t), 2M+1) with int_a^b f(x) cos(t x) dx ~= w @ f(x_j),
    x_j = a + j h, h = (b-a)/(2M) (Filon's rule)."""
    h = (b - a) / (2 * M)
    x = a + h * np.arange(2 * M + 1)
    al, be, ga = _filon_coeffs(t * h)
    cosm = np.cos(np.outer(t, x))
    w = np.empty_like(cosm)
    w[:, 0::2] = be[:, None] * cosm[:, 0::2]
    w[:, 1::2] = ga[:, None] * cosm[:, 1::2]
    w[:, 0] = 0.5 * be * cosm[:, 0] - al * np.sin(t * a)
    w[:, -1] = 0.5 * be * cosm[:, -1] + al * np.sin(t * b)
    return h * w, x


def derivative_tables(kedges: np.ndarray, pairs: List[Tuple[int, float]], nv: int, v_length: float,
                      M: int = 32) -> np.ndarray:
    """(ntab, nk, nv) derivative tables for the (R_fwhm, dv) pairs."""
    v = (v_length / (nv - 1.0)) * np.arange(nv)
    R = np.array([SPEED_OF_LIGHT / r / ONE_SIGMA_2_FWHM for r, _ in pairs])
    dv = np.array([d for _, d in pairs])
    nk = len(kedges) - 1
    out = np.empty((len(pairs), nk, nv))
    for kn in range(nk):
        w, x = filon_cos_weights(kedges[kn], kedges[kn + 1], v, M)
        f = window_sq(x[None, :], dv[:, None], R[:, None]) / MY_PI    # (ntab, 2M+1)
        out[:, kn, :] = f @ w.T
    return out


def signal_tables(binning: Binning, pairs: List[Tuple[int, float]], nv: int, v_length: float,
                  nz: int, pd13=PD13_DEFAULT) -> np.ndarray:
    """(ntab, nv, nz) signal tables on the reference's FFT grid
    (core/sq_table.cpp:221-227,296-311)."""
    itp_dv = v_length / (nv - 1.0)
    nrft = int(2.0 ** (math.ceil(math.log2(2 * 2 * nv)) + 1))
    dxf = itp_dv / 2
    length = dxf * nrft
    k = np.arange(nrft // 2 + 1) * (2 * MY_PI / length)
    zc0 = binning.zbin_centers()[0]
    z1 = zc0 - binning.z_bin_width
    zlen = binning.z_bin_width * (binning.nz + 1)
    zs = z1 + (zlen / (nz - 1.0)) * np.arange(nz)
    out = np.empty((len(pairs), nv, nz))
    beta = pd13[4]
    for t0 in range(0, len(pairs), 64):
        sub = pairs[t0:t0 + 64]
        R = np.array([SPEED_OF_LIGHT / r / ONE_SIGMA_2_FWHM for r, _ in sub])
        dv = np.array([d for _, d in sub])
        w2 = window_sq(k[None, :], dv[:, None], R[:, None])
        if beta == 0.0:
            base = np.fft.irfft(w2 * pd13_power(k, 3.0, *pd13)[None, :], nrft, axis=1) * nrft / length
            samp = base[:, 2 * np.arange(nv)]
            xB = np.power((1.0 + zs) / 4.0, pd13[3])
            out[t0:t0 + len(sub)] = samp[:, :, None] * xB[None, None, :]
        else:
            for iz, z in enumerate(zs):
                fx = np.fft.irfft(w2 * pd13_power(k, z, *pd13)[None, :], nrft, axis=1) * nrft / length
                out[t0:t0 + len(sub), :, iz] = fx[:, 2 * np.arange(nv)]
    return out


@dataclass
class Workload:
    name: str
    binning: Binning
    nv: int
    v_length: float
    nz_table: int
    pd13: tuple
    chunks: List[dict] = field(default_factory=list)      # wave, delta, ivar, R
    pairs: List[Tuple[int, float]] = field(default_factory=list)
    deriv: Optional[np.ndarray] = None
    signal: Optional[np.ndarray] = None
    description: str = ""

    def table_index(self) -> Dict[Tuple[int, float], int]:
        return {p: i for i, p in enumerate(self.pairs)}


def _median_dv(wave):
    d = np.sort(np.log(wave[1:] / wave[:-1]) * SPEED_OF_LIGHT)
    n = d.size
    return (d[n // 2 - 1] + d[n // 2]) / 2 if n % 2 == 0 else d[n // 2]


def _c_round1(x):
    return math.floor(x * 10 + 0.5) / 10 if x >= 0 else -math.floor(-x * 10 + 0.5) / 10


def _realise_delta(rng, vpix, z, R_kms, dv_pix, pd13):
    """Gaussian field with power P_fid W^2 on a fine periodic grid, sampled at vpix."""
    span = vpix[-1] - vpix[0]
    dxf = max(min(dv_pix / 4.0, 2.0), 0.25)
    n = 1 << int(math.ceil(math.log2(2.0 * span / dxf + 16)))
    L = n * dxf
    k = np.arange(n // 2 + 1) * (2 * MY_PI / L)
    p = pd13_power(k, z, *pd13) * window_sq(k, dv_pix, R_kms)
    p[0] = 0.0
    c = (rng.standard_normal(k.size) + 1j * rng.standard_normal(k.size)) * np.sqrt(p / (2 * L))
    c[-1] = c[-1].real * np.sqrt(2.0)
    field = np.fft.irfft(c, n) * n
    grid = dxf * np.arange(n)
    return np.interp(vpix - vpix[0] + 0.25 * L, grid, field)


def _noise_and_mask(rng, npix, median, logsd, frac, runs=None):
    sigma = median * np.exp(logsd * rng.standard_normal(npix))
    ivar = 1.0 / sigma ** 2
    mask = np.zeros(npix, bool)
    if runs is None:
        nm = int(round(frac * npix))
        if nm > 0 and npix > 2:
            mask[1 + rng.choice(npix - 2, size=min(nm, npix - 2), replace=False)] = True
    else:
        target = int(round(frac * npix))
        while mask.sum() < target:
            ln = int(rng.integers(runs[0], runs[1] + 1))
            s = int(rng.integers(1, max(2, npix - ln - 1)))
            mask[s:min(npix - 1, s + ln)] = True
    ivar[mask] = 0.0
    return sigma, ivar


def _finish_tables(w: Workload, with_tables=True):
    pairs = sorted({(c["R"], c["dv_table"]) for c in w.chunks})
    w.pairs = pairs
    if with_tables:
        w.deriv = derivative_tables(w.binning.kband_edges(), pairs, w.nv, w.v_length)
        w.signal = signal_tables(w.binning, pairs, w.nv, w.v_length, w.nz_table, w.pd13)
    return w


def cpu_reference_config(seed: int = 0, nchunks: int = 64, with_tables=True) -> Workload:
    """64 chunks of N=256 at dv=60 km/s (uniform in ln lambda), z=2.4, one
    z bin, Nk=20, sigma_v=70 km/s, log-normal noise, 5% masked, amplitude and
    slope continuum templates in ln lambda."""
    b = Binning(k0=0.0, nklin=10, dklin=0.002, nklog=10, dklog=math.log10(0.05 / 0.02) / 10,
                z0=2.4, z_bin_width=0.2, nz=1, method=1, cont_logl_order=1, cont_lam_order=-1,
                max_forest_length_v=2.0e4)
    w = Workload("cpu-64x256", b, nv=1601, v_length=16000.0, nz_table=9, pd13=PD13_DEFAULT,
                 description="64 chunks, N=256, dv=60 km/s, z=2.4, 1 z bin, Nk=20")
    rng = np.random.default_rng(seed)
    R_fwhm = int(round(SPEED_OF_LIGHT / (70.0 * ONE_SIGMA_2_FWHM)))
    R_kms = SPEED_OF_LIGHT / R_fwhm / ONE_SIGMA_2_FWHM
    lc = LYA_REST * 3.4
    N, dvp = 256, 60.0
    for _ in range(nchunks):
        v = dvp * (np.arange(N) - (N - 1) / 2.0)
        wave = lc * np.exp(v / SPEED_OF_LIGHT)
        delta = _realise_delta(rng, v, 2.4, R_kms, dvp, w.pd13)
        sigma, ivar = _noise_and_mask(rng, N, 0.3, 0.4, 0.05)
        delta = delta + sigma * rng.standard_normal(N)
        w.chunks.append(dict(wave=wave, delta=delta, ivar=ivar, R=R_fwhm,
                             dv_table=_c_round1(_median_dv(wave))))
    return _finish_tables(w, with_tables)


def desi_config(seed: int = 1, nchunks: int = 4096, with_tables=True) -> Workload:
    """DESI-like batch: N ~ U[150,700] on 0.8 A pixels, R ~ U[2000,5000]
    (integer hundreds), z_c ~ U[2.0,4.4]; 13 triangular z bins of width 0.2;
    Nk = 35 (20 linear to 0.02 s/km, 15 log to 0.1 s/km): T = 455."""
    b = Binning(k0=0.0, nklin=20, dklin=0.001, nklog=15, dklog=math.log10(5.0) / 15,
                z0=2.0, z_bin_width=0.2, nz=13, method=1, cont_logl_order=1, cont_lam_order=-1,
                max_forest_length_v=1.2e4)
    w = Workload("desi-4096", b, nv=4601, v_length=46000.0, nz_table=29, pd13=PD13_DEFAULT,
                 description="N~U[150,700] on 0.8 A pixels, R~U[2000,5000], z~U[2.0,4.4], 13 z bins, Nk=35")
    rng = np.random.default_rng(seed)
    zlo, zhi = b.z_lower_edge, b.z_upper_edge
    for _ in range(nchunks):
        N = int(rng.integers(150, 701))
        R_fwhm = int(round(rng.uniform(2000, 5000) / 100.0)) * 100
        zc = rng.uniform(2.0, 4.4)
        lc = LYA_REST * (1 + zc)
        wave = lc + 0.8 * (np.arange(N) - (N - 1) / 2.0)
        keep = (wave >= LYA_REST * (1 + zlo)) & (wave <= LYA_REST * (1 + zhi))
        R_kms = SPEED_OF_LIGHT / R_fwhm / ONE_SIGMA_2_FWHM
        vpix = SPEED_OF_LIGHT * np.log(wave / wave[0])
        dvp = SPEED_OF_LIGHT * 0.8 / lc
        delta = _realise_delta(rng, vpix, zc, R_kms, dvp, w.pd13)
        sigma, ivar = _noise_and_mask(rng, N, 0.3, 0.4, 0.05)
        delta = delta + sigma * rng.standard_normal(N)
        wave, delta, ivar = wave[keep], delta[keep], ivar[keep]
        if wave.size < 100:
            continue
        ivar[0] = ivar[0] or 1.0 / 0.09
        ivar[-1] = ivar[-1] or 1.0 / 0.09
        w.chunks.append(dict(wave=wave, delta=delta, ivar=ivar, R=R_fwhm,
                             dv_table=_c_round1(_median_dv(wave))))
    return _finish_tables(w, with_tables)


def hires_config(seed: int = 2, nchunks: int = 512, with_tables=True, n_range=(800, 2000)) -> Workload:
    """High-resolution batch: N ~ U[800,2000], dv ~ U[2.5,6] km/s (uniform in
    ln lambda), R ~ U[40000,70000] (integer thousands), 13 z bins, Nk = 50 to
    k = 0.3 s/km, noise median 0.1, 10% masked in runs of 5-40 pixels.
    n_range narrows the pixel-count range (tests)."""
    b = Binning(k0=0.0, nklin=30, dklin=0.002, nklog=20, dklog=math.log10(0.3 / 0.06) / 20,
                z0=2.0, z_bin_width=0.2, nz=13, method=1, cont_logl_order=1, cont_lam_order=-1,
                # chunks span 2000-12000 km/s; the short-chunk cut is 0.45 x this
                max_forest_length_v=4.0e3, max_pixel_length_v=150.0)
    w = Workload("hires-512", b, nv=24201, v_length=12100.0, nz_table=29, pd13=PD13_DEFAULT,
                 description="N~U[800,2000], dv~U[2.5,6], R~U[40000,70000], 13 z bins, Nk=50")
    rng = np.random.default_rng(seed)
    for _ in range(nchunks):
        N = int(rng.integers(n_range[0], n_range[1] + 1))
        dvp = round(rng.uniform(2.5, 6.0), 1)
        R_fwhm = int(round(rng.uniform(40000, 70000) / 1000.0)) * 1000
        zc = rng.uniform(2.2, 4.2)
        v = dvp * (np.arange(N) - (N - 1) / 2.0)
        wave = LYA_REST * (1 + zc) * np.exp(v / SPEED_OF_LIGHT)
        R_kms = SPEED_OF_LIGHT / R_fwhm / ONE_SIGMA_2_FWHM
        delta = _realise_delta(rng, v, zc, R_kms, dvp, w.pd13)
        sigma, ivar = _noise_and_mask(rng, N, 0.1, 0.4, 0.10, runs=(5, 40))
        delta = delta + sigma * rng.standard_normal(N)
        w.chunks.append(dict(wave=wave, delta=delta, ivar=ivar, R=R_fwhm,
                             dv_table=_c_round1(_median_dv(wave))))
    return _finish_tables(w, with_tables)


def _gaussian_rows(wave, sigma, nelem, oversampling):
    """Row-normalised Gaussian resolution rows: oversampling == 1 gives the
    nelem-diagonal matrix in mxhelp::DiaMatrix storage (nelem x N: diagonal d
    holds R[c - (nelem//2 - d)][c] at column c), otherwise the
    mxhelp::OversampledMatrix rows (N x nelem) on a grid of spacing
    dlambda / oversampling."""
    N = wave.size
    dlam = np.median(np.diff(wave))
    if oversampling == 1:
        h = nelem // 2
        dense = np.zeros((N, N))
        for i in range(N):
            lo, hi = max(0, i - h), min(N, i + h + 1)
            g = np.exp(-0.5 * ((wave[lo:hi] - wave[i]) / sigma[i]) ** 2)
            dense[i, lo:hi] = g / g.sum()
        vals = np.zeros((nelem, N))
        for d in range(nelem):
            off = h - d
            for c in range(N):
                r = c - off
                if 0 <= r < N:
                    vals[d, c] = dense[r, c]
        return vals
    noff = (nelem - 1) // (2 * oversampling)
    x = (np.arange(nelem) / oversampling - noff) * dlam
    g = np.exp(-0.5 * (x[None, :] / sigma[:, None]) ** 2)
    return g / g.sum(axis=1, keepdims=True)


def rmat_config(seed: int = 3, nchunks: int = 8192, with_tables=True, n_range=(100, 1200),
                kind: str = "oversampled") -> Workload:
    """Resolution-matrix batch: N ~ log-U[100,1200] on 0.8 A pixels, an
    11-diagonal Gaussian resolution matrix per chunk with the per-pixel width
    jittered by +-15% (R ~ U[2000,5000]), oversampled x3 (31 elements per
    row) or kept diagonal (kind "dia"), z ~ U[2.0,4.4], 13 z bins, Nk = 35,
    degree-2 polynomial in ln lambda for the continuum.  Q_k is the analytic
    band window on the matrix's grid; S_fid comes from table 0."""
    b = Binning(k0=0.0, nklin=20, dklin=0.001, nklog=15, dklog=math.log10(5.0) / 15,
                z0=2.0, z_bin_width=0.2, nz=13, method=1, cont_logl_order=2, cont_lam_order=-1,
                # N = 100 pixels at z = 4.4 span 3600 km/s; the short-chunk cut is 0.45 x this
                max_forest_length_v=7.0e3)
    w = Workload(f"rmat-{kind}-{nchunks}", b, nv=8001, v_length=80000.0, nz_table=29, pd13=PD13_DEFAULT,
                 description=f"N~logU[{n_range[0]},{n_range[1]}] on 0.8 A pixels, 11-diagonal Gaussian R "
                             f"({kind}), 13 z bins, Nk=35, 3 templates")
    rng = np.random.default_rng(seed)
    os_ = 3 if kind == "oversampled" else 1
    nelem = 2 * 5 * os_ + 1
    for _ in range(nchunks):
        N = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        R_fwhm = int(round(rng.uniform(2000, 5000) / 100.0)) * 100
        zc = rng.uniform(2.0, 4.4)
        lc = LYA_REST * (1 + zc)
        wave = lc + 0.8 * (np.arange(N) - (N - 1) / 2.0)
        v = SPEED_OF_LIGHT * np.log(wave / lc)
        dvp = float(np.median(np.diff(v)))
        R_kms = SPEED_OF_LIGHT / R_fwhm / ONE_SIGMA_2_FWHM
        delta = _realise_delta(rng, v, zc, R_kms, dvp, w.pd13)
        sigma, ivar = _noise_and_mask(rng, N, 0.3, 0.4, 0.05)
        delta = delta + sigma * rng.standard_normal(N)
        sig_l = wave / (R_fwhm * ONE_SIGMA_2_FWHM) * (1 + rng.uniform(-0.15, 0.15, N))
        # the reader's cutBoundary to the z-bin range, before R is formed
        keep = (wave >= LYA_REST * (1 + b.z_lower_edge)) & (wave <= LYA_REST * (1 + b.z_upper_edge))
        wave, delta, ivar, sig_l = wave[keep], delta[keep], ivar[keep], sig_l[keep]
        if wave.size < n_range[0]:
            continue
        w.chunks.append(dict(wave=wave, delta=delta, ivar=ivar, R=R_fwhm, dv_table=50.0,
                             rmat=dict(kind=kind, nelem=nelem, oversampling=os_,
                                       values=_gaussian_rows(wave, sig_l, nelem, os_))))
    # table 0 supplies S_fid on (z, v); its derivative rows are not used
    w.pairs = [(3000, 50.0)]
    if with_tables:
        w.deriv = derivative_tables(b.kband_edges(), w.pairs, w.nv, w.v_length)
        w.signal = signal_tables(b, w.pairs, w.nv, w.v_length, w.nz_table, w.pd13)
    return w


CONFIGS = {"cpu": cpu_reference_config, "desi": desi_config, "hires": hires_config, "rmat": rmat_config}


def load_engine(engine, w: Workload, chunk_ids: Optional[List[int]] = None):
    """Registers w's tables and chunks with a QMLEngine; returns the list of
    (workload chunk index, engine chunk id) that were accepted."""
    for i, (R, dv) in enumerate(w.pairs):
        engine.add_table(R, dv, w.deriv[i], w.signal[i], w.v_length)
    ok = []
    ids = range(len(w.chunks)) if chunk_ids is None else chunk_ids
    from .qml import QMLError
    for i in ids:
        c = w.chunks[i]
        try:
            if "rmat" in c:
                rm = c["rmat"]
                cid, _ = engine.add_chunk_rmat(c["wave"], c["delta"], c["ivar"], c["R"], rm["kind"],
                                               rm["values"], rm["nelem"], rm["oversampling"])
            else:
                cid, _ = engine.add_chunk(c["wave"], c["delta"], c["ivar"], c["R"])
            ok.append((i, cid))
        except QMLError as e:
            if e.code in (-3, -4, -5, -6):   # the reference's constructor rejections
                continue
            raise
    return ok
