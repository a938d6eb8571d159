"""The reference chunk estimator, restated on the CPU (numpy + LAPACK).

Follows core/chunk_estimate.cpp step by step:

* construction: QSOFile subset copy with masked-boundary cut and median dv
  (io/qso_file.cpp:61-89,130-157,300-329), short-chunk test, table lookup
  (core/sq_table.cpp:388-397), KBIN_UPP (core/chunk_estimate.cpp:32-39,216),
  redshift bins and the triangular extension (:218-257), stored Q list
  (:150-157);
* ``one_iteration``: v/z matrices (:384-398), Q_k from the 1D table
  (:429-480) copied into the other z bins (:871-875), S_fid from the 2D table
  (:400-426), redshift weights (:483-536), C = S + sum theta_k Q_k, the LU
  inverse of N^-1 C + I and the column scaling (:668-719), SVD + sequential
  Sherman-Morrison continuum marginalisation (:576-665), y = C^-1 delta by
  dsymv on the upper triangle, d_k, W_k = C^-1 Q_k by dgemm, b_k, t_k and the
  pairwise Fisher ddot loop (:738-835), accumulation (:893-903).

LAPACK/BLAS: dgetrf+dgetri (LAPACKE_InvertMatrixLU,
mathtools/matrix_helper.cpp:184-204) and dgesvd with jobu='O'
(mathtools/matrix_helper.cpp:437-447) come from scipy's LAPACK; dgemm from
numpy.  Row-major CBLAS "Upper" calls are expressed as products with the
symmetric matrix rebuilt from the upper triangle.

Pins (tests/test_oracle_golden.py): the golden 488x488 S_fid and Q_0 matrices
(all entries), the golden 10x10 Fisher matrix, the golden d/b/t sums and
theta/sigma of tests/truth/test_it1_*; LU inverse against
truth_inverse_matrix.txt; the marginalised inverse against the block-Woodbury
closed form and C_marg^-1 v = 0; the special case S=0, theta=0, no
marginalisation where F = Tr(N^-1 Q N^-1 Q'), b = Tr(N^-1 Q), t = 0; and
brute force in extended precision for N <= 16.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

import numpy as np
from scipy.linalg import lapack

from .bins import (Bins, LYA_REST, MY_PI, SPEED_OF_LIGHT, DOUBLE_EPSILON,
                   TRIANGULAR, redshift_weight)
from .interp import evaluate_1d, evaluate_2d


class ChunkError(RuntimeError):
    """Construction failure; the caller skips the chunk
    (core/one_qso_estimate.cpp:115-125)."""


class LapackError(RuntimeError):
    """LU failure inside an iteration; accumulation is skipped but the chunk
    is kept (core/chunk_estimate.cpp:905-913)."""


@dataclass
class TableSet:
    """SQLookupTable's interpolation grids for one (R, dv) pair
    (core/sq_table.cpp:148-175,325-368)."""
    R: int
    dv: float
    nv: int
    v_length: float
    nz: int
    deriv: np.ndarray   # (nk, nv)
    signal: np.ndarray  # (nv, nz), flat index iz + nz*iv

    def grids(self, bins: Bins):
        v1 = 0.0
        dv = 0.0 if self.nv == 1 else self.v_length / (self.nv - 1.0)
        z1 = bins.zbin_centers[0] - bins.z_bin_width
        zlen = bins.z_bin_width * (bins.nz + 1)
        dz = 0.0 if self.nz == 1 else zlen / (self.nz - 1.0)
        return v1, dv, z1, dz


@dataclass
class ResoMatrix:
    """A chunk's resolution matrix in the reference's storage
    (mathtools/matrix_helper.hpp:122-216).

    kind "dia": mxhelp::DiaMatrix, ``values`` (nelem, size): diagonal d holds
    R[c - (nelem//2 - d)][c] at column c (mathtools/matrix_helper.cpp:504-519,
    770-785).  kind "oversampled": mxhelp::OversampledMatrix, ``values``
    (size, nelem): row i holds R[i][i*oversampling + a], with ncols =
    size*oversampling + nelem - 1 (mathtools/matrix_helper.cpp:833-842)."""
    kind: str
    values: np.ndarray
    nelem: int
    oversampling: int = 1

    def subset(self, i1: int, i2: int) -> "ResoMatrix":
        # Resolution(rmaster, i1, i2), mathtools/matrix_helper.cpp:936-970
        if self.kind == "dia":
            return ResoMatrix("dia", self.values[:, i1:i2].copy(), self.nelem)
        return ResoMatrix("oversampled", self.values[i1:i2].copy(), self.nelem, self.oversampling)

    def dia_row(self, i: int) -> np.ndarray:
        """DiaMatrix::getRow (mathtools/matrix_helper.cpp:608-638): the nelem
        values of row i centred on the diagonal; slots off the matrix take
        the row's own value mirrored about the centre."""
        n, h = self.values.shape[1], self.nelem // 2
        row = np.empty(self.nelem)
        for p in range(self.nelem):
            q = p if 0 <= i - h + p < n else self.nelem - 1 - p
            row[p] = self.values[self.nelem - 1 - q, i - h + q]
        return row

    def deconvolve(self, m: float) -> "ResoMatrix":
        """DiaMatrix::deconvolve (mathtools/matrix_helper.cpp:653-682) with
        RealField::deconvolveSinc (mathtools/real_field.cpp:80-111): each row,
        zero-padded by 5 on both sides, divided by sinc(pi m j / n) in Fourier
        mode j >= 1 (n = nelem + 10; the FFT pair is normalised to identity)."""
        assert self.kind == "dia"
        vals = self.values.copy()
        n, h, pad = vals.shape[1], self.nelem // 2, 5
        nf = self.nelem + 2 * pad
        j = np.arange(1, nf // 2 + 1)
        a = MY_PI * m * j / nf
        for i in range(n):
            x = np.zeros(nf)
            x[pad:pad + self.nelem] = self.dia_row(i)
            X = np.fft.rfft(x)
            X[1:] /= np.sin(a) / a
            y = np.fft.irfft(X, nf)
            for p in range(self.nelem):
                if 0 <= i - h + p < n:
                    vals[self.nelem - 1 - p, i - h + p] = y[pad + p]
        return ResoMatrix("dia", vals, self.nelem)

    def oversample(self, osamp: int, dlambda: float = 1.0) -> "ResoMatrix":
        """Resolution::oversample (mathtools/matrix_helper.cpp:1029-1073):
        per row, shift = min(row) - (smallest magnitude above 1e-15), natural
        cubic spline of log(row - shift) on (i - noff) dlambda, evaluated at
        (jj / osamp - noff) dlambda, shift + exp, normalised to unit sum."""
        assert self.kind == "dia"
        n, noff = self.values.shape[1], self.nelem // 2
        nel = 2 * noff * osamp + 1
        xin = (np.arange(self.nelem) - noff) * dlambda
        xout = (np.arange(nel) / osamp - noff) * dlambda
        rows = np.empty((n, nel))
        for i in range(n):
            row = self.dia_row(i)
            mags = np.abs(row)
            nz = mags[mags > 1e-15]
            shift = row.min() - (nz.min() if nz.size else 1e-15)
            y = np.log(row - shift)
            out = shift + np.exp(natural_cubic_spline(xin, y, xout))
            rows[i] = out * (1.0 / out.sum())
        return ResoMatrix("oversampled", rows, nel, osamp)

    def sandwich(self, fine: np.ndarray, n: int) -> np.ndarray:
        """R fine R^T for a symmetric fine matrix, upper triangle mirrored.
        Oversampled storage follows OversampledMatrix::sandwich
        (mathtools/matrix_helper.cpp:877-902) row by row: v = sum over the
        row's nelem weights of the fine rows from i*os, columns from i*os
        (dgemv, Trans), then B[i][i+j] = v[j*os : j*os+nelem] . R_{i+j}
        (ddot) for j >= 0.  Diagonal storage uses the dense n x n matrix
        (DiaMatrix::sandwich multiplies left, then right-transposed,
        :815-821)."""
        if self.kind != "oversampled":
            r = self.dense(n)
            out = r @ fine @ r.T
        else:
            vals, os_, nel = self.values, self.oversampling, self.nelem
            out = np.zeros((n, n))
            for i in range(n):
                c0 = i * os_
                v = vals[i] @ fine[c0:c0 + nel, c0:]
                win = np.lib.stride_tricks.sliding_window_view(v, nel)[::os_][:n - i]
                out[i, i:] = np.einsum("jt,jt->j", win, vals[i:n])
        iu = np.triu_indices(n, 1)
        out.T[iu] = out[iu]
        return out

    def dense(self, n: int) -> np.ndarray:
        """R as an n x ncols array."""
        if self.kind == "dia":
            r = np.zeros((n, n))
            h = self.nelem // 2
            for d in range(self.nelem):
                off = h - d
                for c in range(n):
                    i = c - off
                    if 0 <= i < n:
                        r[i, c] = self.values[d, c]
            return r
        ncols = n * self.oversampling + self.nelem - 1
        r = np.zeros((n, ncols))
        for i in range(n):
            r[i, i * self.oversampling:i * self.oversampling + self.nelem] = self.values[i]
        return r


def natural_cubic_spline(x: np.ndarray, y: np.ndarray, xe: np.ndarray) -> np.ndarray:
    """Natural cubic spline through (x, y) on a uniform grid, evaluated at xe
    (gsl_interp_cspline, used by Resolution::oversample): second derivatives
    from the tridiagonal system with zero end values."""
    n = x.size
    h = x[1] - x[0]
    m2 = np.zeros(n)
    if n > 2:
        A = np.diag(np.full(n - 2, 4.0)) + np.diag(np.ones(n - 3), 1) + np.diag(np.ones(n - 3), -1)
        r = 6.0 * (y[2:] - 2.0 * y[1:-1] + y[:-2]) / (h * h)
        m2[1:-1] = np.linalg.solve(A, r)
    k = np.clip(np.floor((xe - x[0]) / h).astype(int), 0, n - 2)
    t = (xe - x[k]) / h
    u = 1.0 - t
    return u * y[k] + t * y[k + 1] + h * h / 6.0 * ((u ** 3 - u) * m2[k] + (t ** 3 - t) * m2[k + 1])


def median_sorted(a: np.ndarray) -> float:
    # stats::medianOfSortedArray, mathtools/stats.cpp:8-16
    n = a.size
    if n % 2 == 0:
        return (a[n // 2 - 1] + a[n // 2]) / 2
    return a[n // 2]


def find_spec_res_index(tables: List[TableSet], R: int, dv: float) -> int:
    # core/sq_table.cpp:388-397
    # C round(): halves away from zero
    rounded = np.floor(dv * 10 + 0.5) / 10 if dv >= 0 else -np.floor(-dv * 10 + 0.5) / 10
    for r, t in enumerate(tables):
        if t.R == R and abs(rounded - t.dv) < 0.01:
            return r
    return -1


class Chunk:
    """Mirror of class Chunk (core/chunk_estimate.hpp:46-119)."""

    def __init__(self, bins: Bins, tables: List[TableSet], wave, delta, ivar,
                 R_fwhm: int, i1: int = 0, i2: Optional[int] = None,
                 rmat: Optional[ResoMatrix] = None, exact: bool = False):
        # exact=True departs from the reference's rounding on purpose: v and z
        # come from extended-precision logarithms and square roots (rounded
        # to double once), and the continuum templates are orthonormalised by
        # Gram-Schmidt in extended precision instead of dgesvd.  The tests use
        # it only to measure how far the reference's own rounding moves a
        # case's results (its floor); it is never the parity target itself.
        self.exact = exact
        wave = np.asarray(wave, dtype=np.float64)
        i2 = wave.size if i2 is None else i2
        w = wave[i1:i2].copy()
        d = np.asarray(delta, dtype=np.float64)[i1:i2].copy()
        iv = np.asarray(ivar, dtype=np.float64)[i1:i2].copy()
        # QSOFile::_cutMaskedBoundary, io/qso_file.cpp:139-157
        nz_idx = np.flatnonzero(iv != 0)
        if nz_idx.size == 0:
            raise ChunkError("Empty spectrum when masked pixels at boundary are removed!")
        a, b = nz_idx[0], nz_idx[-1] + 1
        w, d, iv = w[a:b], d[a:b], iv[a:b]
        # the resolution matrix follows the cut (io/qso_file.cpp:82-85)
        self.rmat = rmat.subset(i1 + a, i1 + b) if rmat is not None else None
        self.bins = bins
        self.N = w.size
        self.real_size = int(np.count_nonzero(iv))
        # recalcDvDLam, io/qso_file.cpp:325-329 (median of the sorted array)
        if self.N > 1:
            self.dv_kms = median_sorted(np.sort(np.log(w[1:] / w[:-1]) * SPEED_OF_LIGHT))
        else:
            self.dv_kms = -1.0
        self.R_fwhm = int(R_fwhm)
        # _copyQSOFile short-chunk test, core/chunk_estimate.cpp:172-187
        lenv = SPEED_OF_LIGHT * np.log(w[-1] / w[0]) if self.N > 0 else 0.0
        minnopix = int(lenv / bins.max_pixel_length_v)
        if self.real_size < minnopix or lenv < bins.min_forest_length_v:
            raise ChunkError("Short chunk")
        # with a resolution matrix the table index is 0 (core/chunk_estimate.cpp:189-193)
        self.res_index = 0 if self.rmat is not None else find_spec_res_index(tables, self.R_fwhm, self.dv_kms)
        if self.res_index == -1 or not tables:
            raise ChunkError(f"SpecRes R: {self.R_fwhm}, dv: {self.dv_kms} not found in tables!")
        self.table = tables[self.res_index]
        # readMinMaxMedRedshift, io/qso_file.cpp:280-300
        self.lower_z = w[0] / LYA_REST - 1
        self.upper_z = w[-1] / LYA_REST - 1
        self.median_z = median_sorted(w) / LYA_REST - 1
        self.kbin_upp = bins.max_k_index(MY_PI / self.dv_kms)
        # _findRedshiftBin, core/chunk_estimate.cpp:225-236
        self.zbin = bins.find_redshift_bin(self.median_z)
        self.zbin_low = bins.find_redshift_bin(self.lower_z)
        self.zbin_upp = bins.find_redshift_bin(self.upper_z)
        self.wave = w / LYA_REST
        self.delta = d
        self.ivar = iv
        if self.rmat is not None:
            # fine grid (core/chunk_estimate.cpp:936-944): oversampled R spans
            # ncols points of spacing dlambda / LYA_REST / oversampling, the
            # chunk's median dlambda (io/qso_file.cpp:21-29,325-329)
            if self.rmat.kind == "oversampled":
                dlam = median_sorted(np.sort(w[1:] - w[:-1]))
                fine = dlam / LYA_REST / self.rmat.oversampling
                ncols = self.N * self.rmat.oversampling + self.rmat.nelem - 1
                disp = self.rmat.nelem // 2
                self.fine_wave = self.wave[0] + (np.arange(ncols) - disp) * fine
            else:
                self.fine_wave = self.wave.copy()
            self.R = self.rmat.dense(self.N)
        self._set_nq_and_fisher_index()
        nk = bins.nk
        self.stored = [i for i in range(self.n_q) if (i + self.fisher_index_start) % nk < self.kbin_upp]
        self.dbt = np.zeros((3, self.n_q))
        self.fisher = np.zeros((self.n_q, self.n_q))

    def _set_nq_and_fisher_index(self):
        # core/chunk_estimate.cpp:240-257
        b = self.bins
        nq = self.zbin_upp - self.zbin_low + 1
        fis = b.nk * self.zbin_low
        if b.method == TRIANGULAR:
            if self.lower_z < b.zbin_centers[self.zbin_low] and self.zbin_low != 0:
                nq += 1
                fis -= b.nk
            if b.zbin_centers[self.zbin_upp] < self.upper_z and self.zbin_upp < b.nz - 1:
                nq += 1
        self.n_q = nq * b.nk
        self.fisher_index_start = fis

    # -- matrices -----------------------------------------------------------
    def vz_matrices(self):
        """_setVZMatrices, core/chunk_estimate.cpp:384-398, mirrored."""
        return _vz_upper(self.wave, self.exact)

    def _mirror(self, m):
        iu = np.triu_indices(self.N, 1)
        m.T[iu] = m[iu]
        return m

    def _fine_vz(self):
        """v/z on the fine grid of the resolution matrix (upper, mirrored),
        computed once per chunk."""
        if getattr(self, "_fine_cache", None) is None:
            v, z = _vz_upper(self.fine_wave, self.exact)
            self._fine_cache = (v, z, np.triu_indices(self.fine_wave.size))
        return self._fine_cache

    def _sandwich(self, fine):
        """Resolution::sandwich (mathtools/matrix_helper.cpp:1075-1084)."""
        return self.rmat.sandwich(fine, self.N)

    def signal_matrix(self, v, z):
        """_setFiducialSignalMatrix, core/chunk_estimate.cpp:400-426."""
        v1, dv, z1, dz = self.table.grids(self.bins)
        if self.rmat is not None:
            fv, fz, iu = self._fine_vz()
            n = fv.shape[0]
            s = np.zeros((n, n))
            s[iu] = evaluate_2d(fz[iu], fv[iu], z1, dz, v1, dv, self.table.signal.reshape(-1),
                                self.table.nz, self.table.nv)
            s.T[np.triu_indices(n, 1)] = s[np.triu_indices(n, 1)]
            return self._sandwich(s)
        iu = np.triu_indices(self.N)
        s = np.zeros((self.N, self.N))
        s[iu] = evaluate_2d(z[iu], v[iu], z1, dz, v1, dv, self.table.signal.reshape(-1),
                            self.table.nz, self.table.nv)
        return self._mirror(s)

    def q_base(self, v, kn):
        """_setQiMatrix for a first-z-bin k band, core/chunk_estimate.cpp:429-480."""
        if self.rmat is not None:
            # analytic band window on the fine grid, then R . Q . R^T (:444-454,471-472)
            kc = self.bins.kband_centers[kn]
            dk = self.bins.kband_edges[kn + 1] - self.bins.kband_edges[kn]
            fv, _, iu = self._fine_vz()
            n = fv.shape[0]
            vv = fv[iu]
            with np.errstate(invalid="ignore", divide="ignore"):
                x = dk * vv / 2
                val = np.where(vv == 0, dk / MY_PI, (dk / MY_PI) * np.cos(kc * vv) * (np.sin(x) / x))
            q = np.zeros((n, n))
            q[iu] = val
            q.T[np.triu_indices(n, 1)] = q[np.triu_indices(n, 1)]
            return self._sandwich(q)
        v1, dv, _, _ = self.table.grids(self.bins)
        iu = np.triu_indices(self.N)
        q = np.zeros((self.N, self.N))
        q[iu] = evaluate_1d(v[iu], v1, dv, self.table.deriv[kn])
        return self._mirror(q)

    def redshift_weights(self, z, zm):
        """The weight matrix built row-wise in _applyRedshiftInterp
        (core/chunk_estimate.cpp:483-536): for j >= i it is w_zm(z_ij)."""
        iu = np.triu_indices(self.N)
        t = np.zeros((self.N, self.N))
        t[iu] = redshift_weight(self.bins, z[iu], zm)
        return self._mirror(t)

    def growth_factors(self, z, zm):
        """_scaleDerivativesWithRedshiftGrowth, core/chunk_estimate.cpp:539-573:
        ((1 + z_ij) / (1 + z_m))^B for j >= i, mirrored.  The reference
        scales all N^2 entries with the z matrix it holds; away from the
        oversampled path only its upper triangle is set (:384-398), so the
        symmetric form here is the reference's upper triangle."""
        iu = np.triu_indices(self.N)
        g = np.zeros((self.N, self.N))
        zcm = self.bins.zbin_centers[zm] + 1.0
        g[iu] = np.power((z[iu] + 1.0) / zcm, self.bins.fiducial_redshift_power)
        return self._mirror(g)

    def q_matrices(self, v, z):
        """Stored Q_k after copies and z weights, in stored order."""
        nk = self.bins.nk
        base = {}
        for kn in range(self.kbin_upp):
            base[kn] = self.q_base(v, kn)
        nzq = len(self.stored) // self.kbin_upp if self.kbin_upp else 0
        out = []
        for iz in range(nzq):
            zm = self.fisher_index_start // nk + iz
            w = self.redshift_weights(z, zm)
            if self.bins.redshift_growth_on:
                w = w * self.growth_factors(z, zm)
            for s in self.stored[iz * self.kbin_upp:(iz + 1) * self.kbin_upp]:
                out.append(base[s % nk] * w)
        return out

    # -- the iteration -------------------------------------------------------
    def inverse_covariance(self, sfid, qs, theta, mode="lu"):
        """setInvertCovarianceMatrix, core/chunk_estimate.cpp:668-719."""
        N = self.N
        c = np.zeros((N, N)) if self.bins.turn_off_sfid else sfid.copy()
        for i_kz, q in zip(self.stored, qs):
            c += theta[self.fisher_index_start + i_kz] * q
        iv = self.ivar
        if mode == "lu":
            m = c * iv[:, None]
            m[np.diag_indices(N)] += 1.0
            lu, piv, info = lapack.dgetrf(m)
            if info != 0:
                raise LapackError("ERROR in LU decomposition.")
            inv, info = lapack.dgetri(lu, piv)
            if info != 0:
                raise LapackError("ERROR in LU invert.")
            cinv = inv * iv[None, :]
        else:
            # SPD form N^{-1/2} (I + N^{-1/2} S N^{-1/2})^{-1} N^{-1/2}: the
            # same matrix, used only to measure the rounding floor
            ds = np.sqrt(iv)
            a = c * ds[:, None] * ds[None, :]
            a[np.diag_indices(N)] += 1.0
            l, info = lapack.dpotrf(a, lower=1)
            if info != 0:
                raise LapackError("ERROR in Cholesky decomposition.")
            ainv, info = lapack.dpotri(l, lower=1)
            ainv = np.tril(ainv) + np.tril(ainv, -1).T
            cinv = ainv * ds[:, None] * ds[None, :]
        if self.bins.cont_nvecs > 0:
            cinv = self.add_marginalizations(cinv)
        return cinv

    def templates(self):
        """_addMarginalizations template rows, core/chunk_estimate.cpp:576-594,
        609-623 (wave in units of lambda_Lya, masked pixels included)."""
        b, N, w = self.bins, self.N, self.wave
        rows = [np.full(N, 1.0 / np.sqrt(N))]
        for cmo in range(1, b.cont_logl_order + 1):
            t = np.power(np.log(w), cmo)
            rows.append(t / np.linalg.norm(t))
        for cmo in range(1, b.cont_lam_order + 1):
            t = np.power(w, cmo)
            rows.append(t / np.linalg.norm(t))
        return np.array(rows)

    def add_marginalizations(self, cinv):
        """SVD (jobu='O') + sequential Sherman-Morrison,
        core/chunk_estimate.cpp:596-665."""
        vt = self.templates()                      # (nvecs, N) == col-major N x nvecs
        u, s, _vt, info = lapack.dgesvd(vt.T.copy(order="F"), compute_uv=1, full_matrices=0)
        if info != 0:
            raise LapackError("ERROR in SVD.")
        if self.exact:
            u = _orthonormal_extended(vt.T)
        nuse = len(s)
        while nuse > 0:
            if s[nuse - 1] / s[0] > DOUBLE_EPSILON:
                break
            nuse -= 1
        cinv = cinv.copy()
        for i in range(nuse):
            v = u[:, i]
            y = _symv_upper(cinv, v)
            norm = -1.0 / np.dot(v, y)
            cinv += norm * np.outer(y, y)
        return cinv

    def ps_before_fisher(self, cinv, sfid, qs, times=None):
        """computePSbeforeFvector, core/chunk_estimate.cpp:738-835."""
        N = self.N
        dk, nk_, tk = self.dbt
        y = _symv_upper(cinv, self.delta)
        ws = []
        t0 = time.perf_counter()
        for i_kz, q in zip(self.stored, qs):
            dk[i_kz] = np.dot(y, _symv_upper(q, y))
            ws.append(cinv @ q)
        if times is not None:
            times["set_modqs"] = times.get("set_modqs", 0.0) + time.perf_counter() - t0
        iv = self.ivar
        with np.errstate(divide="ignore"):
            inv_iv = np.where(iv != 0, 1.0 / np.where(iv != 0, iv, 1.0), 0.0)
        ncinv = cinv * inv_iv[:, None]
        for i_kz, wmat in zip(self.stored, ws):
            nk_[i_kz] = np.dot(wmat.ravel(), ncinv.ravel())
        if not self.bins.turn_off_sfid:
            scinv = sfid @ cinv
            for i_kz, wmat in zip(self.stored, ws):
                tk[i_kz] = np.dot(wmat.ravel(), scinv.ravel())
        if self.bins.use_precomputed_fisher:
            return ws
        t0 = time.perf_counter()
        for a, (i_kz, wa) in enumerate(zip(self.stored, ws)):
            wt = wa.T.ravel()
            for j_kz, wb in zip(self.stored[a:], ws[a:]):
                self.fisher[i_kz, j_kz] = np.dot(wt, wb.ravel())
        if times is not None:
            times["set_fisher"] = times.get("set_fisher", 0.0) + time.perf_counter() - t0
        return ws

    def one_iteration(self, theta, fisher_sum, dbt_sum, mode="lu", times=None):
        """Chunk::oneQSOiteration, core/chunk_estimate.cpp:837-914.  times: a
        dict accumulating seconds per phase in the reference's mytime
        categories (core/global_numbers.cpp:476-502): set_qs (v/z, Q_k and
        their z weights), set_sfid, c_inv (the inverse with the continuum
        marginalisation), set_modqs (d_k and W_k), set_fisher, and dbt_rest
        (the b/t contractions)."""
        clock = time.perf_counter
        t0 = clock()
        v, z = self.vz_matrices()
        qs = self.q_matrices(v, z)
        t1 = clock()
        sfid = None if self.bins.turn_off_sfid else self.signal_matrix(v, z)
        t2 = clock()
        try:
            cinv = self.inverse_covariance(sfid, qs, theta, mode)
            t3 = clock()
            inner = {} if times is not None else None
            self.ps_before_fisher(cinv, sfid, qs, inner)
            t4 = clock()
        except LapackError:
            return False
        if times is not None:
            for k, dt in (("set_qs", t1 - t0), ("set_sfid", t2 - t1), ("c_inv", t3 - t2)):
                times[k] = times.get(k, 0.0) + dt
            for k, dt in inner.items():
                times[k] = times.get(k, 0.0) + dt
            times["dbt_rest"] = times.get("dbt_rest", 0.0) + (t4 - t3) - sum(inner.values())
        accumulate(self, fisher_sum, dbt_sum)
        return True


def _vz_upper(wave, exact):
    """v_ij = c ln(l_j / l_i) and z_ij = sqrt(l_i l_j) - 1 for j >= i (zero
    below), in double as the reference, or from extended precision."""
    n = wave.size
    iu = np.triu_indices(n)
    w = wave.astype(np.longdouble) if exact else wave
    c = np.longdouble(SPEED_OF_LIGHT) if exact else SPEED_OF_LIGHT
    v = np.zeros((n, n))
    z = np.zeros((n, n))
    v[iu] = (c * np.log(w[None, :] / w[:, None]))[iu]
    z[iu] = (np.sqrt(w[:, None] * w[None, :]) - 1)[iu]
    return v, z


def _orthonormal_extended(a):
    """Orthonormal basis of the columns of a by modified Gram-Schmidt, twice,
    in extended precision (for the rounding-floor measurement only)."""
    q = a.astype(np.longdouble).copy()
    for _ in range(2):
        for j in range(q.shape[1]):
            for i in range(j):
                q[:, j] -= (q[:, i] @ q[:, j]) * q[:, i]
            q[:, j] /= np.sqrt(q[:, j] @ q[:, j])
    return q.astype(np.float64)


def _symv_upper(a, x):
    """cblas_dsymv(RowMajor, Upper): only a[i, j] with j >= i is read."""
    au = np.triu(a)
    return au @ x + np.triu(a, 1).T @ x


def accumulate(ch: Chunk, fisher_sum: np.ndarray, dbt_sum: np.ndarray) -> None:
    """core/chunk_estimate.cpp:893-903: upper triangle at offset
    (T+1)*fisher_index_start, and d/b/t at fisher_index_start."""
    fis, nq = ch.fisher_index_start, ch.n_q
    iu = np.triu_indices(nq)
    blk = fisher_sum[fis:fis + nq, fis:fis + nq]
    blk[iu] += ch.fisher[iu]
    dbt_sum[:, fis:fis + nq] += ch.dbt


def decide_indices(wave: np.ndarray, bins: Bins) -> List[int]:
    """OneQSOEstimate::decideIndices, core/one_qso_estimate.cpp:17-50."""
    size = wave.size
    vel = SPEED_OF_LIGHT * np.log(wave / wave[0])
    vmax = vel[-1]
    nchunks = 1
    if bins.number_of_chunks > 1:
        nchunks += int(vmax / bins.max_forest_length_per_chunk)
        vmax = min(bins.max_forest_length_per_chunk, vmax / nchunks)
    idx = [0]
    lo = 0
    vel = vel.copy()
    for _ in range(nchunks):
        jj = lo + int(np.searchsorted(vel[lo:], vmax, side="right"))
        idx.append(jj)
        if jj == size:
            break
        lo = jj
        vel[lo:] -= vel[lo]
    return idx
