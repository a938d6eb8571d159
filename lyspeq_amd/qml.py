"""ctypes binding of ``libqmlhip.so`` (include/lyspeq_qml.h).

``QMLEngine`` is the Python face of the C ABI; ``Chunk`` mirrors the
reference's class of the same name (core/chunk_estimate.hpp:46-119): its
constructor registers a chunk and raises the reference's errors, and
``QMLEngine.iterate`` runs what every Chunk::oneQSOiteration of one
Newton-Raphson iteration does, accumulating into the caller's fisher_sum and
dbt_sum.  There is no CPU fallback: without the built library, or without a
GPU, every call raises.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass
from typing import List, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libqmlhip.so")

PREC_FP64, PREC_FP32, PREC_BF16X3, PREC_BF16X4 = 0, 1, 2, 3
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3, "bf16x4": PREC_BF16X4}

QML_OK = 0
QML_ERR_ARG, QML_ERR_HIP, QML_ERR_SHORT_CHUNK, QML_ERR_NO_TABLE = -1, -2, -3, -4
QML_ERR_EMPTY, QML_ERR_MEMORY, QML_ERR_LIMIT = -5, -6, -7


class QMLError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class ShortChunkError(QMLError):
    """std::runtime_error("Short chunk ...") of core/chunk_estimate.cpp:172-187."""


class TableNotFoundError(QMLError):
    """std::out_of_range of core/chunk_estimate.cpp:198-206."""


class EmptyChunkError(QMLError):
    """io/qso_file.cpp:150-153."""


class ChunkMemoryError(QMLError):
    """The chunk's matrices do not fit the device scratch budget
    (Chunk::_setStoredMatrices, core/chunk_estimate.cpp:258-282)."""


_ERRORS = {QML_ERR_SHORT_CHUNK: ShortChunkError, QML_ERR_NO_TABLE: TableNotFoundError,
           QML_ERR_EMPTY: EmptyChunkError, QML_ERR_MEMORY: ChunkMemoryError}


class _Opts(C.Structure):
    _fields_ = [("device", C.c_int), ("precision", C.c_int), ("scratch_gb", C.c_double),
                ("max_batch_chunks", C.c_int)]


class _Bins(C.Structure):
    _fields_ = [("nk", C.c_int), ("nz", C.c_int), ("kband_edges", C.POINTER(C.c_double)),
                ("zbin_centers", C.POINTER(C.c_double)), ("z_bin_width", C.c_double),
                ("z_binning_method", C.c_int), ("cont_logl_order", C.c_int),
                ("cont_lam_order", C.c_int), ("turn_off_sfid", C.c_int),
                ("max_forest_length_v", C.c_double), ("max_pixel_length_v", C.c_double),
                ("redshift_growth_on", C.c_int), ("fiducial_redshift_power", C.c_double)]


class _ChunkIn(C.Structure):
    _fields_ = [("wave", C.POINTER(C.c_double)), ("delta", C.POINTER(C.c_double)),
                ("ivar", C.POINTER(C.c_double)), ("size", C.c_int), ("R_fwhm", C.c_int)]


class _RMat(C.Structure):
    _fields_ = [("kind", C.c_int), ("nelem", C.c_int), ("oversampling", C.c_int),
                ("values", C.POINTER(C.c_double))]


RMAT_KINDS = {"dia": 1, "oversampled": 2}


class ChunkInfo(C.Structure):
    _fields_ = [("size", C.c_int), ("real_size", C.c_int), ("res_index", C.c_int),
                ("kbin_upp", C.c_int), ("zbin", C.c_int), ("zbin_low", C.c_int),
                ("zbin_upp", C.c_int), ("n_q_matrices", C.c_int),
                ("fisher_index_start", C.c_int), ("n_stored", C.c_int),
                ("n_templates", C.c_int), ("dv_kms", C.c_double), ("lower_z", C.c_double),
                ("upper_z", C.c_double), ("median_z", C.c_double)]


class Timings(C.Structure):
    _fields_ = [("assemble_ms", C.c_double), ("invert_ms", C.c_double),
                ("weight_ms", C.c_double), ("fisher_ms", C.c_double),
                ("accumulate_ms", C.c_double), ("total_ms", C.c_double),
                ("n_batches", C.c_longlong), ("kernel_ms", C.c_double * 13)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_ if f != "kernel_ms"}
        d["kernel_ms"] = {name: self.kernel_ms[i] for i, name in enumerate(KERNEL_NAMES)}
        return d


KERNEL_NAMES = ["assemble", "cholesky", "trinv", "solve", "dpass", "small", "xgen",
                "whiten_left", "whiten_right", "gram", "finalize", "accumulate", "rmat"]


_lib = None

EXPORTS = ["qml_create", "qml_destroy", "qml_last_error", "qml_set_bins", "qml_add_table",
           "qml_add_chunk", "qml_add_chunk_rmat", "qml_commit", "qml_iterate", "qml_get_chunk_result",
           "qml_get_timings", "qml_set_pipeline", "qml_set_chunk_active", "qml_num_chunks",
           "qml_total_kz_bins", "qml_rmat_deconvolve", "qml_rmat_oversample", "qml_make_tables",
           "qml_set_graph", "qml_autotune", "qml_bootstrap", "qml_poisson_draws", "qml_set_kernel_stages",
           "qml_get_schedule", "qml_bootstrap_complete", "qml_bootstrap_median", "qml_bootstrap_median_noalloc"]


EXPORTS += ["qml_set_stack_slab", "qml_get_batch", "qml_set_np_buckets", "qml_chunk_matrix",
            "qml_chunk_marginalize"]

# qml_chunk_matrix's matrices (Chunk's protected builders)
CMAT_V, CMAT_Z, CMAT_SFID, CMAT_Q, CMAT_ZWEIGHT, CMAT_GROWTH = range(6)


class _Schedule(C.Structure):
    _fields_ = [("max_batch_chunks", C.c_int), ("n_batches", C.c_int), ("stages_left", C.c_int),
                ("stages_gram", C.c_int), ("pipeline", C.c_int), ("graph", C.c_int), ("stack_slab", C.c_int),
                ("max_chunk_gb", C.c_double), ("stages_right", C.c_int), ("np_buckets", C.c_int)]


class _BatchInfo(C.Structure):
    _fields_ = [("n_chunks", C.c_int), ("np_min", C.c_int), ("np_max", C.c_int), ("members_max", C.c_int)]


class _TableSpec(C.Structure):
    _fields_ = [("nv", C.c_int), ("v_length", C.c_double), ("nz", C.c_int), ("pd13", C.c_double * 6),
                ("filon_panels", C.c_int)]


def load_library(path: str = LIB_PATH):
    """Loads the HIP library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m lyspeq_amd.build`")
    lib = C.CDLL(path)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    lib.qml_create.argtypes = [C.POINTER(vp), C.POINTER(_Opts)]
    lib.qml_destroy.argtypes = [vp]
    lib.qml_last_error.restype = C.c_char_p
    lib.qml_set_bins.argtypes = [vp, C.POINTER(_Bins)]
    lib.qml_add_table.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.c_double, C.c_int, dp, dp,
                                  C.POINTER(C.c_int)]
    lib.qml_add_chunk.argtypes = [vp, C.POINTER(_ChunkIn), C.POINTER(ChunkInfo), C.POINTER(C.c_int)]
    lib.qml_rmat_deconvolve.argtypes = [C.c_int, C.c_int, C.c_double, dp]
    lib.qml_rmat_oversample.argtypes = [C.c_int, C.c_int, dp, C.c_int, dp]
    lib.qml_make_tables.argtypes = [vp, C.POINTER(_TableSpec), C.c_int, C.POINTER(C.c_int), dp, dp, dp]
    lib.qml_add_chunk_rmat.argtypes = [vp, C.POINTER(_ChunkIn), C.POINTER(_RMat), C.POINTER(ChunkInfo),
                                       C.POINTER(C.c_int)]
    lib.qml_commit.argtypes = [vp]
    lib.qml_iterate.argtypes = [vp, dp, dp, dp]
    lib.qml_get_chunk_result.argtypes = [vp, C.c_int, dp, dp, C.POINTER(C.c_int)]
    lib.qml_get_timings.argtypes = [vp, C.POINTER(Timings)]
    lib.qml_set_pipeline.argtypes = [vp, C.c_int]
    lib.qml_set_graph.argtypes = [vp, C.c_int]
    lib.qml_set_kernel_stages.argtypes = [vp, C.c_int, C.c_int, C.c_int]
    lib.qml_get_schedule.argtypes = [vp, C.POINTER(_Schedule)]
    lib.qml_set_stack_slab.argtypes = [vp, C.c_int]
    lib.qml_get_batch.argtypes = [vp, C.c_int, C.POINTER(_BatchInfo)]
    lib.qml_set_np_buckets.argtypes = [vp, C.c_int]
    lib.qml_chunk_matrix.argtypes = [vp, C.c_int, C.c_int, C.c_int, dp, C.POINTER(C.c_int)]
    lib.qml_chunk_marginalize.argtypes = [vp, C.c_int, dp]
    lib.qml_bootstrap.argtypes = [vp, C.c_int, C.c_ulonglong, C.c_int, C.POINTER(C.c_int), dp, dp, dp]
    lib.qml_poisson_draws.argtypes = [C.c_ulonglong, C.c_longlong, C.POINTER(C.c_int)]
    ip = C.POINTER(C.c_int)
    lib.qml_bootstrap_complete.argtypes = [vp, C.c_int, C.c_ulonglong, C.c_int, ip, ip, dp, ip, ip, dp, dp]
    lib.qml_bootstrap_median.argtypes = [vp, C.c_int, C.c_int, dp, dp, dp]
    lib.qml_bootstrap_median_noalloc.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp]
    lib.qml_autotune.argtypes = [vp, dp, C.POINTER(C.c_int), C.c_int, C.c_int, C.POINTER(C.c_int), dp]
    lib.qml_set_chunk_active.argtypes = [vp, C.c_int, C.c_int]
    lib.qml_num_chunks.argtypes = [vp]
    lib.qml_total_kz_bins.argtypes = [vp]
    _lib = lib
    return lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(code: int, lib):
    if code != QML_OK:
        msg = lib.qml_last_error().decode()
        raise _ERRORS.get(code, QMLError)(code, msg)


def rmat_deconvolve(values, m: float) -> np.ndarray:
    """qml_rmat_deconvolve (DiaMatrix::deconvolve) on a copy of a whole
    spectrum's diagonal storage, ndiags x size."""
    lib = load_library()
    v = np.array(values, dtype=np.float64, order="C")
    nd, size = v.shape
    code = lib.qml_rmat_deconvolve(size, nd, float(m), _dptr(v))
    if code != QML_OK:
        raise QMLError(code, "qml_rmat_deconvolve: invalid arguments")
    return v


def poisson_draws(seed: int, n: int) -> np.ndarray:
    """n unit-mean Poisson draws of the reference's PoissonRNG(seed) (host)."""
    lib = load_library()
    out = np.zeros(int(n), dtype=np.int32)
    _check(lib.qml_poisson_draws(int(seed), int(n), out.ctypes.data_as(C.POINTER(C.c_int))), lib)
    return out


def rmat_oversample(values, oversampling: int) -> np.ndarray:
    """qml_rmat_oversample (Resolution::oversample): a whole spectrum's
    diagonal storage (ndiags x size) to oversampled rows, size x nelem."""
    lib = load_library()
    v = np.ascontiguousarray(values, dtype=np.float64)
    nd, size = v.shape
    out = np.empty((size, 2 * (nd // 2) * int(oversampling) + 1))
    code = lib.qml_rmat_oversample(size, nd, _dptr(v), int(oversampling), _dptr(out))
    if code != QML_OK:
        raise QMLError(code, "qml_rmat_oversample: invalid arguments")
    return out


@dataclass
class Binning:
    """bins:: as set up by bins::setUpBins (core/global_numbers.cpp:209-251)
    plus the specifics:: keys the chunk path reads."""
    k0: float
    nklin: int
    dklin: float
    nklog: int
    dklog: float
    z0: float
    z_bin_width: float
    nz: int
    klast: float = -1.0
    method: int = 1
    cont_logl_order: int = 1
    cont_lam_order: int = -1
    turn_off_sfid: bool = False
    max_forest_length_v: float = 3.5e4
    number_of_chunks: int = 1
    max_pixel_length_v: float = 150.0
    # RedshiftGrowthOn (core/global_numbers.cpp:127) and its exponent B,
    # FiducialRedshiftPower (core/fiducial_cosmology.cpp:204, compiled-in
    # default 3.591244 at :268-271)
    redshift_growth_on: bool = False
    fiducial_redshift_power: float = 3.591244

    def kband_edges(self) -> np.ndarray:
        nk = self.nklin + self.nklog
        ktemp = (self.k0 + self.dklin * self.nklin) * math.pow(10, self.nklog * self.dklog)
        edges = [self.k0 + self.dklin * i for i in range(self.nklin + 1)]
        for i in range(1, self.nklog + 1):
            edges.append(edges[self.nklin] * math.pow(10.0, i * self.dklog))
        if self.klast > ktemp:
            edges.append(self.klast)
        return np.array(edges)

    def kband_centers(self) -> np.ndarray:
        e = self.kband_edges()
        return (e[:-1] + e[1:]) / 2.0

    def zbin_centers(self) -> np.ndarray:
        return np.array([self.z0 + self.z_bin_width * m for m in range(self.nz)])

    @property
    def nk(self) -> int:
        return len(self.kband_edges()) - 1

    @property
    def total_kz_bins(self) -> int:
        return self.nk * self.nz

    @property
    def z_lower_edge(self) -> float:
        return self.zbin_centers()[0] - self.z_bin_width / 2.0

    @property
    def z_upper_edge(self) -> float:
        return self.zbin_centers()[-1] + self.z_bin_width / 2.0


class QMLEngine:
    """One device context holding tables and chunks for repeated iterations."""

    def __init__(self, binning: Binning, precision="fp64", device: int = 0,
                 scratch_gb: float = 0.0, max_batch_chunks: int = 0, stack_slab: int = 0):
        self._lib = load_library()
        prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        opts = _Opts(device, prec, scratch_gb, max_batch_chunks)
        self._ctx = C.c_void_p()
        _check(self._lib.qml_create(C.byref(self._ctx), C.byref(opts)), self._lib)
        self.binning = binning
        self.precision = prec
        self._edges = np.ascontiguousarray(binning.kband_edges(), dtype=np.float64)
        self._zc = np.ascontiguousarray(binning.zbin_centers(), dtype=np.float64)
        b = _Bins(len(self._edges) - 1, binning.nz, _dptr(self._edges), _dptr(self._zc),
                  binning.z_bin_width, binning.method, binning.cont_logl_order,
                  binning.cont_lam_order, int(bool(binning.turn_off_sfid)),
                  binning.max_forest_length_v / max(1, binning.number_of_chunks),
                  binning.max_pixel_length_v, int(bool(binning.redshift_growth_on)),
                  float(binning.fiducial_redshift_power))
        _check(self._lib.qml_set_bins(self._ctx, C.byref(b)), self._lib)
        if stack_slab:
            self.set_stack_slab(stack_slab)
        self.T = len(self._edges) - 1
        self.T *= binning.nz
        self.tables = []
        self.chunk_infos: List[ChunkInfo] = []

    def close(self):
        if getattr(self, "_ctx", None):
            self._lib.qml_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_table(self, R_fwhm: int, dv_kms: float, deriv: np.ndarray, signal: Optional[np.ndarray],
                  v_length: float) -> int:
        deriv = np.ascontiguousarray(deriv, dtype=np.float64)
        nk, nv = deriv.shape
        if nk != self.binning.nk:
            raise ValueError("derivative table has the wrong number of k bands")
        if signal is not None:
            signal = np.ascontiguousarray(signal, dtype=np.float64)
            nz = signal.shape[1]
            sp = _dptr(signal)
        else:
            nz, sp = 2, None
        r = C.c_int(-1)
        _check(self._lib.qml_add_table(self._ctx, int(R_fwhm), float(dv_kms), nv, float(v_length), nz,
                                       _dptr(deriv), sp, C.byref(r)), self._lib)
        self.tables.append((int(R_fwhm), float(dv_kms)))
        return r.value

    def make_tables(self, pairs, nv: int, v_length: float, nz: int, pd13, filon_panels: int = 0,
                    deriv: bool = True, signal: bool = True):
        """qml_make_tables: the S/Q tables of the (R_fwhm, dv) pairs on the
        device, as (deriv (npairs, nk, nv), signal (npairs, nv, nz)); an
        output not asked for is None."""
        n = len(pairs)
        R = np.ascontiguousarray([int(r) for r, _ in pairs], dtype=np.int32)
        dv = np.ascontiguousarray([float(d) for _, d in pairs], dtype=np.float64)
        q = np.empty((n, self.binning.nk, nv)) if deriv else None
        s = np.empty((n, nv, nz)) if signal else None
        spec = _TableSpec(int(nv), float(v_length), int(nz), (C.c_double * 6)(*[float(x) for x in pd13]),
                          int(filon_panels))
        _check(self._lib.qml_make_tables(self._ctx, C.byref(spec), n, R.ctypes.data_as(C.POINTER(C.c_int)),
                                         _dptr(dv), _dptr(q) if deriv else None, _dptr(s) if signal else None),
               self._lib)
        return q, s

    def add_chunk(self, wave, delta, ivar, R_fwhm: int):
        w = np.ascontiguousarray(wave, dtype=np.float64)
        d = np.ascontiguousarray(delta, dtype=np.float64)
        iv = np.ascontiguousarray(ivar, dtype=np.float64)
        cin = _ChunkIn(_dptr(w), _dptr(d), _dptr(iv), w.size, int(R_fwhm))
        info = ChunkInfo()
        cid = C.c_int(-1)
        _check(self._lib.qml_add_chunk(self._ctx, C.byref(cin), C.byref(info), C.byref(cid)), self._lib)
        self.chunk_infos.append(info)
        return cid.value, info

    def add_chunk_rmat(self, wave, delta, ivar, R_fwhm: int, kind: str, values, nelem: int,
                       oversampling: int = 1):
        """add_chunk with a resolution matrix in the reference's storage:
        kind "dia" (values nelem x size) or "oversampled" (size x nelem)."""
        w = np.ascontiguousarray(wave, dtype=np.float64)
        d = np.ascontiguousarray(delta, dtype=np.float64)
        iv = np.ascontiguousarray(ivar, dtype=np.float64)
        vals = np.ascontiguousarray(values, dtype=np.float64)
        cin = _ChunkIn(_dptr(w), _dptr(d), _dptr(iv), w.size, int(R_fwhm))
        rm = _RMat(RMAT_KINDS[kind], int(nelem), int(oversampling), _dptr(vals))
        info = ChunkInfo()
        cid = C.c_int(-1)
        _check(self._lib.qml_add_chunk_rmat(self._ctx, C.byref(cin), C.byref(rm), C.byref(info), C.byref(cid)),
               self._lib)
        self.chunk_infos.append(info)
        return cid.value, info

    def commit(self):
        _check(self._lib.qml_commit(self._ctx), self._lib)

    def iterate(self, theta: np.ndarray, fisher_sum: np.ndarray, dbt_sum: np.ndarray):
        """Adds every chunk's F (upper) and d, b, t into the given arrays."""
        th = np.ascontiguousarray(theta, dtype=np.float64)
        assert th.size == self.T and fisher_sum.shape == (self.T, self.T) and dbt_sum.shape == (3, self.T)
        assert fisher_sum.flags.c_contiguous and dbt_sum.flags.c_contiguous
        _check(self._lib.qml_iterate(self._ctx, _dptr(th), _dptr(fisher_sum), _dptr(dbt_sum)), self._lib)

    def chunk_result(self, cid: int):
        nq = self.chunk_infos[cid].n_q_matrices
        F = np.zeros((nq, nq))
        dbt = np.zeros((3, nq))
        st = C.c_int(0)
        _check(self._lib.qml_get_chunk_result(self._ctx, cid, _dptr(F), _dptr(dbt), C.byref(st)), self._lib)
        return F, dbt, st.value

    def chunk_matrix(self, cid: int, which: int, arg: int = 0) -> np.ndarray:
        """One matrix of the reference Chunk's protected builders for chunk
        cid (qml_chunk_matrix): which is one of the CMAT_* values, arg the
        chunk-local i_kz (CMAT_Q) or the z bin (CMAT_ZWEIGHT, CMAT_GROWTH)."""
        n = C.c_int(0)
        _check(self._lib.qml_chunk_matrix(self._ctx, cid, int(which), int(arg), None, C.byref(n)), self._lib)
        out = np.empty((n.value, n.value))
        _check(self._lib.qml_chunk_matrix(self._ctx, cid, int(which), int(arg), _dptr(out), C.byref(n)), self._lib)
        return out

    def chunk_marginalize(self, cid: int, cinv: np.ndarray) -> np.ndarray:
        """Chunk::_addMarginalizations on a copy of cinv (qml_chunk_marginalize)."""
        out = np.ascontiguousarray(cinv, dtype=np.float64).copy()
        _check(self._lib.qml_chunk_marginalize(self._ctx, cid, _dptr(out)), self._lib)
        return out

    def set_pipeline(self, on: bool):
        """Stage overlap across two streams (default on); off gives per-kernel
        durations free of contention, with identical results."""
        _check(self._lib.qml_set_pipeline(self._ctx, int(bool(on))), self._lib)

    def set_graph(self, on: bool):
        """Graph replay (default off): the next iteration captures its launch
        sequence into a HIP graph, later iterations replay it with the new
        theta.  Results are identical; a replay reports total_ms only."""
        _check(self._lib.qml_set_graph(self._ctx, int(bool(on))), self._lib)

    def autotune(self, theta: np.ndarray, caps=(64, 128, 256, 512), reps: int = 1):
        """Plans the batches again under each cap on chunks per batch, times
        reps iterations (after one warm-up) at theta and keeps the fastest
        schedule.  Returns (best cap, its ms per iteration)."""
        th = np.ascontiguousarray(theta, dtype=np.float64)
        assert th.size == self.T
        cv = (C.c_int * len(caps))(*[int(x) for x in caps])
        best = C.c_int(0)
        ms = C.c_double(0.0)
        _check(self._lib.qml_autotune(self._ctx, _dptr(th), cv, len(caps), int(reps), C.byref(best),
                                      C.byref(ms)), self._lib)
        return best.value, ms.value

    def bootstrap(self, nboots: int, qso_of_chunk, nqso: int, seed: int = 0, realizations: bool = True):
        """FastBootstrap of the last iteration's per-chunk results: returns
        (realizations nboots x T or None, mean T, covariance T x T)."""
        q = np.ascontiguousarray(qso_of_chunk, dtype=np.int32)
        assert q.size == len(self.chunk_infos)
        R = np.zeros((nboots, self.T)) if realizations else None
        mean, cov = np.zeros(self.T), np.zeros((self.T, self.T))
        _check(self._lib.qml_bootstrap(self._ctx, int(nboots), int(seed), int(nqso),
                                       q.ctypes.data_as(C.POINTER(C.c_int)), _dptr(R) if R is not None else None,
                                       _dptr(mean), _dptr(cov)), self._lib)
        return R, mean, cov

    def bootstrap_complete(self, nboots: int, qso_of_chunk, nqso: int, seed: int = 0, weights=None):
        """The bootstrap's complete form (a Cholesky solve per realisation):
        returns (valid realisations nvalid x T, valid flags, mean, cov)."""
        ip = C.POINTER(C.c_int)
        q = np.ascontiguousarray(qso_of_chunk, dtype=np.int32)
        assert q.size == len(self.chunk_infos)
        w = None
        if weights is not None:
            w = np.ascontiguousarray(weights, dtype=np.int32)
            assert w.shape == (nboots, nqso)
        R = np.zeros((nboots, self.T))
        valid = np.zeros(nboots, dtype=np.int32)
        nv = C.c_int(0)
        mean, cov = np.zeros(self.T), np.zeros((self.T, self.T))
        _check(self._lib.qml_bootstrap_complete(self._ctx, int(nboots), int(seed), int(nqso), q.ctypes.data_as(ip),
                                                w.ctypes.data_as(ip) if w is not None else None, _dptr(R),
                                                valid.ctypes.data_as(ip), C.byref(nv), _dptr(mean), _dptr(cov)),
               self._lib)
        return R[:nv.value], valid.astype(bool), mean, cov

    def bootstrap_median(self, realizations):
        """The median bootstrap of realisations (nboots x T): returns (median
        T, median covariance T x T).  Needs no chunks or commit."""
        R = np.ascontiguousarray(realizations, dtype=np.float64)
        nboots, T = R.shape
        med, mad = np.zeros(T), np.zeros((T, T))
        _check(self._lib.qml_bootstrap_median(self._ctx, int(nboots), int(T), _dptr(R), _dptr(med), _dptr(mad)),
               self._lib)
        return med, mad

    def set_kernel_stages(self, left: int, right: int, gram: int):
        """Operand staging of the split-BF16 left, right and Gram products:
        1, 2 through registers with one or two LDS stages (2 not for the
        right product), 3, 4 direct-to-LDS loads with two or three stages."""
        _check(self._lib.qml_set_kernel_stages(self._ctx, int(left), int(right), int(gram)), self._lib)

    def set_stack_slab(self, members: int):
        """Streams each chunk's stack in slabs of `members` (0: all at once)."""
        _check(self._lib.qml_set_stack_slab(self._ctx, int(members)), self._lib)

    def schedule(self) -> dict:
        s = _Schedule()
        _check(self._lib.qml_get_schedule(self._ctx, C.byref(s)), self._lib)
        return {k: getattr(s, k) for k, _ in s._fields_}

    def set_np_buckets(self, on: bool):
        """Also end a batch where the padded size changes (re-plans if committed)."""
        _check(self._lib.qml_set_np_buckets(self._ctx, int(bool(on))), self._lib)

    def batches(self) -> list:
        """The current plan's batches in launch order: chunk count, padded
        size range (the Np bucket edges) and largest member count."""
        out = []
        for b in range(self.schedule()["n_batches"]):
            info = _BatchInfo()
            _check(self._lib.qml_get_batch(self._ctx, b, C.byref(info)), self._lib)
            out.append({k: getattr(info, k) for k, _ in info._fields_})
        return out

    def set_chunk_active(self, cid: int, active: bool):
        """Leaves a chunk out of (or back in) the accumulated sums."""
        _check(self._lib.qml_set_chunk_active(self._ctx, int(cid), int(bool(active))), self._lib)

    def timings(self) -> dict:
        t = Timings()
        _check(self._lib.qml_get_timings(self._ctx, C.byref(t)), self._lib)
        return t.as_dict()


class Chunk:
    """Chunk(qmaster, i1, i2) of core/chunk_estimate.cpp:141-167: registers
    pixels [i1, i2) of a spectrum with the engine.  Raises ShortChunkError,
    TableNotFoundError or EmptyChunkError like the reference constructor."""

    def __init__(self, engine: QMLEngine, wave, delta, ivar, R_fwhm: int, i1: int = 0,
                 i2: Optional[int] = None):
        i2 = len(wave) if i2 is None else i2
        self.engine = engine
        self.id, info = engine.add_chunk(np.asarray(wave)[i1:i2], np.asarray(delta)[i1:i2],
                                         np.asarray(ivar)[i1:i2], R_fwhm)
        self.info = info
        self.fisher_index_start = info.fisher_index_start
        self.N_Q_MATRICES = info.n_q_matrices
        self.KBIN_UPP = info.kbin_upp
        self.ZBIN, self.ZBIN_LOW, self.ZBIN_UPP = info.zbin, info.zbin_low, info.zbin_upp

    def size(self) -> int:
        return self.info.size

    def realSize(self) -> int:
        return self.info.real_size

    @property
    def fisher_matrix(self) -> np.ndarray:
        return self.engine.chunk_result(self.id)[0]

    @property
    def dbt_estimate_before_fisher_vector(self) -> np.ndarray:
        return self.engine.chunk_result(self.id)[1]
