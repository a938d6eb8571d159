"""Builds the HIP library ``lyspeq_amd/libqmlhip.so`` for gfx950 in-tree.

Usage: ``python -m lyspeq_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libqmlhip.so")

SOURCES = ["kernels_chunk.hip", "kernels_factor.hip", "kernels_gemm.hip", "kernels_rmat.hip", "kernels_tables.hip",
           "kernels_boot.hip", "qml_api.cpp",
           "rmat_prep.cpp"]

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC",
    # no packed-FP32 instructions in any generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        cmd = [hipcc(), *HIPCC_FLAGS, "-I", os.path.join(ROOT, "include"), "-I", CSRC,
               "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr}")
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    build_host(verbose)
    return LIB


HOST_LIB = os.path.join(HERE, "libqmlhost.so")
HOST_DRIVER = os.path.join(ROOT, "tests", "cpp", "host_driver")
DRIVER_CHECK = os.path.join(ROOT, "tests", "cpp", "driver_check")
DROPIN_CHECK = os.path.join(ROOT, "tests", "cpp", "dropin_check")


def _run(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{what} failed:\n{r.stderr}")


def build_host(verbose: bool = False) -> str:
    """The C++ host layer (Chunk / OneQSOEstimate / Engine over the C ABI)
    as libqmlhost.so, and the C++ test driver that exercises it."""
    rocm_lib = "/opt/rocm/lib"
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "host"),
           "-I", os.path.join(HERE, "host", "dropin")]
    link = ["-L", HERE, "-lqmlhip", "-L", rocm_lib, f"-Wl,-rpath-link,{rocm_lib}"]
    tmp = HOST_LIB + ".tmp"
    # the drop-in Chunk / OneQSOEstimate, with this build's mirrors of the
    # reference globals and reader (replaced by the reference's own inside its tree)
    dropin = [os.path.join(HERE, "host", "dropin", f) for f in ("dropin.cpp", "mirror.cpp", "io/picca_file.cpp",
                                                                 "io/bootstrap_file.cpp")]
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", *inc,
          os.path.join(HERE, "host", "qml_host.cpp"), os.path.join(HERE, "host", "qml_driver.cpp"), *dropin, *link, "-lz",
          "-Wl,-rpath,$ORIGIN", "-o", tmp],
         "host library build", verbose)
    os.replace(tmp, HOST_LIB)
    for src, exe in ((os.path.join(ROOT, "tests", "cpp", "host_driver.cpp"), HOST_DRIVER),
                     (os.path.join(ROOT, "tests", "cpp", "driver_check.cpp"), DRIVER_CHECK),
                     (os.path.join(ROOT, "tests", "cpp", "dropin_check.cpp"), DROPIN_CHECK)):
        if os.path.exists(src):
            _run(["g++", "-O2", "-std=c++17", "-Wall", *inc, src, "-L", HERE, "-lqmlhost", *link,
                  "-Wl,-rpath,$ORIGIN/../../lyspeq_amd", "-o", exe], f"{os.path.basename(exe)} build", verbose)
    return HOST_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
