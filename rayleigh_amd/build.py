"""Builds the HIP library rayleigh_amd/libraysht.so for gfx950, in tree.

    python -m rayleigh_amd.build          (or __graft_entry__.build())

Every .hip file under csrc/ is compiled with hipcc for gfx950 only; the small
quad-precision host part (csrc/host_quad.cpp) with g++.  Objects go to
rayleigh_amd/_build/ (git-ignored); the shared library sits next to this file.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libraysht.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# per-file additions: the double-double table builder must not have products fused into its sums
FILE_FLAGS = {"legendre_tables.hip": ["-ffp-contract=off"]}

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    # packed FP32 instructions stay off in every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(os.path.dirname(HERE), "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for src in sorted(glob.glob(os.path.join(CSRC, "*.hip"))):
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        extra = FILE_FLAGS.get(os.path.basename(src), [])
        jobs.append([HIPCC] + HIP_FLAGS + extra + ["-c", src, "-o", obj])
    host = os.path.join(CSRC, "host_quad.cpp")
    jobs.append(["g++", "-O2", "-std=gnu++17", "-fPIC", "-Wall", "-c", host,
                 "-o", os.path.join(BUILD, "host_quad.cpp.o")])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and (r.stdout or r.stderr):
                print(r.stdout + r.stderr, file=sys.stderr)
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lquadmath"])
    os.replace(tmp, LIB)
    build_fortran()
    return LIB


FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_BUILD = os.path.join(FORTRAN_DIR, "_build")
FORTRAN_TEST = os.path.join(FORTRAN_BUILD, "test_sht_raysht")
FORTRAN_RATES = os.path.join(FORTRAN_BUILD, "rates_raysht")
ROCM_LIB = "/opt/rocm/lib"
FC = os.environ.get("RAYSHT_FC", "/opt/rocm/bin/amdflang")


def build_fortran() -> str | None:
    """The ISO_C_BINDING module (RaySHT_Math_Layer), the Fortran test program that drives
    it and the program that times its two call forms; skipped (None) where no Fortran
    compiler is installed."""
    if not os.path.exists(FC):
        return None
    os.makedirs(FORTRAN_BUILD, exist_ok=True)
    srcs = [os.path.join(FORTRAN_DIR, "raysht_structures.F90"), os.path.join(FORTRAN_DIR, "raysht_math_layer.F90")]
    _run_in(FORTRAN_BUILD, [FC, "-O2", "-fPIC", "-c"] + srcs)
    for prog, out in (("test_sht_raysht.F90", FORTRAN_TEST), ("rates_raysht.F90", FORTRAN_RATES)):
        src = os.path.join(os.path.dirname(HERE), "tests", "fortran", prog)
        _run_in(FORTRAN_BUILD, [FC, "-O2", "-o", out, src, "raysht_structures.o", "raysht_math_layer.o",
                                "-L" + HERE, "-lraysht", "-L" + ROCM_LIB, "-lamdhip64",
                                "-Wl,-rpath,$ORIGIN/../..", "-Wl,-rpath," + ROCM_LIB])
    return FORTRAN_TEST


def _run_in(cwd, cmd):
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=cwd)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


if __name__ == "__main__":
    print(build(verbose=True))
