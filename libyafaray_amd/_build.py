"""Compile the HIP library (libymx.so) for gfx950 in-tree.

The library is plain C ABI over hand-written HIP kernels; it is built with
hipcc directly (no torch extension) so any host language can load it.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libymx.so"

SOURCES = [
    "accel_build.hip",
    "accel_trace.hip",
    "accel_api.hip",
    "pt_kernels.hip",
    "pt_api.hip",
    "guide.hip",
]

# -ffp-contract=off / IEEE division / denormals kept: FP32 results must match
# the reference's unfused x86 arithmetic bit for bit.
# -packed-fp32-ops off: packed FP32 instructions are not generated.
HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libymx.so)")


def _compile(src: str) -> Path:
    out = OBJ / (Path(src).stem + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", str(ROOT / "include"), "-c", str(CSRC / src), "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stdout}\n{res.stderr}")
    return out


def build_library(verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), 8)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


HOST_SRC = PKG / "host" / "ymx_host.cpp"
HOST_LIB = PKG / "libymx_host.so"
CPP_TEST_SRC = ROOT / "tests" / "cpp" / "test_host.cpp"
CPP_TEST_BIN = ROOT / "tests" / "cpp" / "bin" / "test_host"

# The host layer keeps the kernels' float arithmetic (single-ray traversal of
# the exported tree), so it is compiled without contraction or fast math too.
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
              "-Wall", "-Wno-unknown-pragmas"]


def build_host(verbose: bool = False) -> Path:
    """The C++ host layer (libymx_host.so) over libymx.so's C ABI."""
    tmp = HOST_LIB.with_suffix(".so.tmp")
    cmd = [shutil.which("g++") or "g++", *HOST_FLAGS, "-shared", "-I", str(ROOT / "include"),
           "-o", str(tmp), str(HOST_SRC), "-L", str(PKG), "-lymx", "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"host layer build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, HOST_LIB)
    if verbose:
        print(f"built {HOST_LIB}")
    return HOST_LIB


def build_cpp_tests(verbose: bool = False) -> Path:
    """The C++ test driver of the host layer (links the oracle as its checker)."""
    CPP_TEST_BIN.parent.mkdir(exist_ok=True)
    tmp = CPP_TEST_BIN.with_suffix(".tmp")
    cmd = [hipcc(), "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
           "-Wno-unused-command-line-argument", "-Wno-unknown-pragmas",
           "-I", str(ROOT / "include"), "-o", str(tmp), str(CPP_TEST_SRC),
           "-L", str(PKG), "-lymx_host", "-lymx", "-L", str(ROOT / "oracle"), "-loracle",
           "-Wl,-rpath,$ORIGIN/../../../libyafaray_amd", "-Wl,-rpath,$ORIGIN/../../../oracle"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"C++ test build failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, CPP_TEST_BIN)
    if verbose:
        print(f"built {CPP_TEST_BIN}")
    return CPP_TEST_BIN


def build_oracle(verbose: bool = False) -> Path:
    """Compile the CPU oracle (test infrastructure) with its own Makefile."""
    res = subprocess.run(["make", "-C", str(ROOT / "oracle"), "-s"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    if verbose:
        print("built oracle/liboracle.so")
    return ROOT / "oracle" / "liboracle.so"


REF_ROOT = Path("/root/reference")


def build_ref(verbose: bool = False) -> Path | None:
    """oracle/_ref/libref.so: the reference's own polygon test, samplers, fast
    math and filter, compiled from its sources (test infrastructure).  Only
    where the reference is mounted; elsewhere the tests that use it skip."""
    if not (REF_ROOT / "include" / "geometry" / "shape" / "shape_polygon.h").exists():
        if verbose:
            print("reference sources not mounted: oracle/_ref not built")
        return None
    res = subprocess.run(["make", "-C", str(ROOT / "oracle"), "-f", "ref.mk", "-s", f"REF={REF_ROOT}"],
                         capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle/_ref build failed:\n{res.stdout}\n{res.stderr}")
    if verbose:
        print("built oracle/_ref/libref.so")
    return ROOT / "oracle" / "_ref" / "libref.so"


if __name__ == "__main__":
    build_library(verbose=True)
    build_host(verbose=True)
    build_oracle(verbose=True)
    build_ref(verbose=True)
    build_cpp_tests(verbose=True)
