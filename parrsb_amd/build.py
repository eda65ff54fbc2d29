"""Builds parrsb_amd/libparrsb_amd.so in-tree: HIP kernels for gfx950 with
hipcc, the C host layer with gcc, linked into one shared library that
exports the C ABI declared in include/."""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
HIP_DIR = os.path.join(PKG, "csrc", "hip")
HOST_DIR = os.path.join(PKG, "csrc", "host")
OUT_DIR = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libparrsb_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")
ARCH = os.environ.get("PARRSB_AMD_ARCH", "gfx950")
MPI_INC = "/opt/conda/include"

HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-I" + INC,
             # no packed FP32 instructions (v_pk_*_f32) in any kernel
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-result", "-Wno-unused-value"]
C_FLAGS = ["-O2", "-fPIC", "-std=c11", "-Wall", "-Wno-unused-function", "-I" + INC]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + p.stdout + p.stderr)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return ""


def sources():
    hip = sorted(os.path.join(HIP_DIR, f) for f in os.listdir(HIP_DIR) if f.endswith(".hip"))
    host = sorted(os.path.join(HOST_DIR, f) for f in os.listdir(HOST_DIR) if f.endswith(".c"))
    return hip, host


def build(verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    hip, host = sources()
    cflags = list(C_FLAGS)
    if os.path.exists(os.path.join(MPI_INC, "mpi.h")):
        cflags.append("-I" + MPI_INC)
    else:
        cflags.append("-DPARRSB_NO_MPI")
    jobs = []
    objs = []
    for src in hip:
        obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", "-o", obj, src])
        objs.append(obj)
    for src in host:
        obj = os.path.join(OUT_DIR, os.path.basename(src) + ".o")
        jobs.append([CC] + cflags + ["-c", "-o", obj, src])
        objs.append(obj)
    nproc = min(len(jobs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(nproc) as ex:
        for f in [ex.submit(_run, j) for j in jobs]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC"] + objs + ["-o", tmp])
    shutil.move(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
