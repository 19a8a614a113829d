"""Build the CPU checker.  TEST INFRASTRUCTURE ONLY.

``oracle/liboracle.so``      -- txoracle.c, the CPU restatement (OpenMP, AVX2+FMA
                                baseline ISA like the reference's configure).
``oracle/_ref/libhough_ref.so`` -- TransientX's own include/hough_transform.h,
                                compiled from where it lies in the reference tree
                                through the harness oracle/ref_hough.cpp.  Built
                                only when the reference tree is present.

Compiler dumps and objects stay out of the repository (.gitignore).
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
REF = Path(os.environ.get("TX_REFERENCE_ROOT", "/root/reference"))


def run(cmd):
    print(" ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


def build_oracle() -> Path:
    out = HERE / "liboracle.so"
    tmp = HERE / "liboracle.so.tmp"
    run(["gcc", "-O3", "-std=c11", "-fPIC", "-shared", "-fopenmp", "-march=x86-64-v3",
         "-ffp-contract=off", "-Wall", "-o", tmp, HERE / "txoracle.c", "-lm"])
    os.replace(tmp, out)
    return out


def build_ref() -> Path | None:
    hdr = REF / "include" / "hough_transform.h"
    if not hdr.exists():
        print("reference tree absent: oracle/_ref not rebuilt", flush=True)
        return None
    outdir = HERE / "_ref"
    outdir.mkdir(exist_ok=True)
    out = outdir / "libhough_ref.so"
    tmp = outdir / "libhough_ref.so.tmp"
    # the reference's own flags: configure.ac sets -std=c++11 -O3 and
    # AX_CHECK_X86_FEATURES adds -mavx2 -mfma on an AVX2/FMA host
    # (config/ax_check_x86_features.m4:76), so GCC contracts a + b*c into FMA
    run(["g++", "-O3", "-std=c++11", "-mavx2", "-mfma", "-fPIC", "-shared", "-I", REF / "include",
         "-o", tmp, HERE / "ref_hough.cpp"])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build_oracle()
    build_ref()
    sys.exit(0)
