"""Build the in-tree native libraries.

``libtxgpu.so``  -- the HIP kernels and C ABI (include/txgpu.h), gfx950 only.
``oracle``       -- the CPU checker, built by oracle/build.py (tests only).

Run ``python -m transientx_amd.build`` or call :func:`build_all`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libtxgpu.so"
ARCH = os.environ.get("TX_OFFLOAD_ARCH", "gfx950")

# Packed FP32 instructions are kept out of generated code on purpose.
HIP_FLAGS = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    "-Wno-unused-value",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, cwd=None):
    print(" ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, cwd=cwd)


def build_txgpu(verbose: bool = True) -> Path:
    hipcc = _hipcc()
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    srcs = [CSRC / "kernels.hip", CSRC / "txgpu.hip", CSRC / "plan.cpp", CSRC / "host.cpp",
            CSRC / "refine.hip", CSRC / "diff.hip", CSRC / "rfi.hip", CSRC / "dump.cpp",
            CSRC / "sigproc.cpp"]
    objs = []
    for src in srcs:
        if not src.exists():
            continue
        obj = objdir / (src.stem + ".o")
        lang = ["-x", "hip"] if src.suffix == ".hip" else []
        _run([hipcc, *HIP_FLAGS, *lang, "-I", ROOT / "include", "-c", src, "-o", obj])
        objs.append(obj)
    tmp = LIB.with_suffix(".so.tmp")
    _run([hipcc, *HIP_FLAGS, "-shared", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


DROPIN_SRC = ROOT / "tests" / "cpp" / "dropin_search.cpp"
DROPIN_BIN = ROOT / "tests" / "cpp" / "dropin_search"


def build_dropin_test() -> Path | None:
    """The C++ drop-in (include/transientx_gpu.hpp) driven like SinglePulse,
    linked against libtxgpu.so: a plain g++ host program."""
    if not DROPIN_SRC.exists():
        return None
    tmp = DROPIN_BIN.with_suffix(".tmp")
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-I", ROOT / "include", DROPIN_SRC, "-o", tmp,
          "-L", PKG, "-ltxgpu", "-Wl,-rpath,$ORIGIN/../../transientx_amd",
          "-Wl,-rpath-link,/opt/rocm/lib"])
    os.replace(tmp, DROPIN_BIN)
    return DROPIN_BIN


CALLS_SRC = ROOT / "tests" / "cpp" / "dropin_calls.cpp"
CALLS_BIN = ROOT / "tests" / "cpp" / "dropin_calls"


def build_calls_test() -> Path | None:
    """Every call expression of the reference's callers on the drop-in, at
    the reference's -std=c++11: a plain g++ host program."""
    if not CALLS_SRC.exists():
        return None
    tmp = CALLS_BIN.with_suffix(".tmp")
    _run(["g++", "-std=c++11", "-O2", "-Wall", "-I", ROOT / "include", CALLS_SRC, "-o", tmp,
          "-L", PKG, "-ltxgpu", "-Wl,-rpath,$ORIGIN/../../transientx_amd",
          "-Wl,-rpath-link,/opt/rocm/lib"])
    os.replace(tmp, CALLS_BIN)
    return CALLS_BIN


def build_all() -> None:
    build_txgpu()
    build_dropin_test()
    build_calls_test()
    # the checker (CPU oracle + reference header harness) for tests and bench
    oracle_build = ROOT / "oracle" / "build.py"
    if oracle_build.exists():
        _run([sys.executable, oracle_build])


if __name__ == "__main__":
    build_all()
