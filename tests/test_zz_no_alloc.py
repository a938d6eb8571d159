"""The allocation-free median bootstrap (qml_bootstrap_median_noalloc)
against qml_bootstrap_median, through the C ABI, bit for bit: two calls in a
row on device arrays, the second smaller (the shared scratch keeps the first
call's contents), then a size beyond the scratch rejected."""
import ctypes as C

import numpy as np
import pytest

from lyspeq_amd import qml, synth
from lyspeq_amd.qml import QMLEngine

pytestmark = pytest.mark.gpu


def _engine():
    return QMLEngine(synth.cpu_reference_config(nchunks=2, with_tables=False).binning, precision="fp64")


class _Hip:
    """Device buffers through the HIP runtime itself (hipMalloc / hipMemcpy /
    hipFree), for the caller's side of the test."""

    def __init__(self):
        self.rt = C.CDLL("libamdhip64.so")
        self.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.rt.hipFree.argtypes = [C.c_void_p]

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0   # host to device
        return p

    def get(self, p, shape):
        out = np.empty(shape, dtype=np.float64)
        assert self.rt.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self, *ps):
        for p in ps:
            assert self.rt.hipFree(p) == 0


def _noalloc(lib, hip, eng, R):
    nboots, T = R.shape
    dR = hip.put(R)
    dmed = hip.put(np.full(T, np.nan))
    dmad = hip.put(np.full((T, T), np.nan))
    rc = lib.qml_bootstrap_median_noalloc(eng._ctx, nboots, T, dR, dmed, dmad)
    med, mad = hip.get(dmed, (T,)), hip.get(dmad, (T, T))
    hip.free(dR, dmed, dmad)
    return rc, med, mad


@pytest.mark.parametrize("seed", [0, 1])
def test_noalloc_median_bootstrap_matches_allocating_entry_point(seed):
    print(f"[case] allocation-free median bootstrap, seed {seed}", flush=True)
    lib = qml.load_library()
    rng = np.random.default_rng(seed)
    eng = _engine()
    hip = _Hip()
    # first call larger than the second: 4000 x 300 and 1500 x 70, with ties
    # (one decimal) in the second
    R1 = rng.standard_normal((4000, 300)) * rng.uniform(0.5, 3.0, 300)
    R2 = np.round(rng.standard_normal((1500, 70)) * rng.uniform(0.5, 3.0, 70), 1)
    for R in (R1, R2):
        rc, med, mad = _noalloc(lib, hip, eng, R)
        assert rc == 0, lib.qml_last_error()
        wmed, wmad = eng.bootstrap_median(R)
        assert np.array_equal(med, wmed)
        assert np.array_equal(mad, wmad)
    # beyond the static scratch: rejected before any launch
    big = np.zeros((20000, 455))
    rc, _, _ = _noalloc(lib, hip, eng, big)
    assert rc == qml.QML_ERR_LIMIT
    assert b"exceeds" in lib.qml_last_error()
    # host memory is rejected, not launched on
    h = np.zeros(16)
    rc = lib.qml_bootstrap_median_noalloc(eng._ctx, 4, 4, h.ctypes.data_as(C.c_void_p),
                                          h.ctypes.data_as(C.c_void_p), h.ctypes.data_as(C.c_void_p))
    assert rc == qml.QML_ERR_ARG
    eng.close()
