"""The C-ABI library loads and exports every symbol include/lyspeq_qml.h
declares (no compute calls: this runs without a GPU)."""
import os
import re

import pytest

from lyspeq_amd import qml

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    with open(os.path.join(ROOT, "include", "lyspeq_qml.h")) as f:
        src = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(qml_\w+)\s*\(", src, re.M)))


def test_header_and_binding_agree():
    assert sorted(qml.EXPORTS) == _declared()


def test_library_exports_every_symbol():
    if not os.path.exists(qml.LIB_PATH):
        pytest.skip("library not built (run `python -m lyspeq_amd.build`)")
    lib = qml.load_library()
    for name in _declared():
        assert hasattr(lib, name), name


def test_no_device_fails_loudly():
    if not os.path.exists(qml.LIB_PATH):
        pytest.skip("library not built")
    import ctypes as C
    lib = qml.load_library()
    ctx = C.c_void_p()
    opts = qml._Opts(0, 0, 0.0, 0)
    rc = lib.qml_create(C.byref(ctx), C.byref(opts))
    if rc == 0:
        lib.qml_destroy(ctx)
        pytest.skip("a device is present")
    assert rc == qml.QML_ERR_HIP
    assert b"no HIP device" in lib.qml_last_error()
