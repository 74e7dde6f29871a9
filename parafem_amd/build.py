"""Build libparafem_amd.so (HIP, gfx950) in-tree.

``python -m parafem_amd.build`` compiles every ``csrc/*.hip`` with hipcc for
gfx950 and links them into ``parafem_amd/libparafem_amd.so``, then builds the
Fortran binding (``fortran/pf_gpu.f90``) and the p121, p122, p123 and p124
drivers (``bin/pf_p12*``) with amdflang.  It always recompiles (the caller decides
when to build).
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libparafem_amd.so")
OBJDIR = os.path.join(HERE, "_build")
FSRC = os.path.join(HERE, "fortran")
P121 = os.path.join(HERE, "bin", "pf_p121")
P122 = os.path.join(HERE, "bin", "pf_p122")
P123 = os.path.join(HERE, "bin", "pf_p123")
P124 = os.path.join(HERE, "bin", "pf_p124")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off (v_pk_*_f32 are never emitted)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FP64 vector arithmetic must round every operation as written (the PCG
    # trajectory is compared bit for bit); fused multiply-adds are explicit.
    "-ffp-contract=off",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def _compile(src: str) -> str:
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build_fortran(verbose: bool = False) -> str | None:
    """Compile the Fortran binding module and the p121, p122 and p124 drivers against the
    library (rpath to this directory).  Skipped, returning None, when no
    Fortran compiler is installed."""
    if not os.path.exists(FLANG):
        return None
    moddir = os.path.join(OBJDIR, "fmod")
    os.makedirs(moddir, exist_ok=True)
    os.makedirs(os.path.dirname(P121), exist_ok=True)
    for exe, src in ((P121, "pf_p121.f90"), (P122, "pf_p122.f90"), (P123, "pf_p123.f90"),
                     (P124, "pf_p124.f90")):
        tmp = exe + ".tmp"
        # no contraction: the host-array loop rounds every product and sum
        # as written, as the device does (fused multiply-adds are explicit)
        cmd = [FLANG, "-O2", "-ffp-contract=off", "-J", moddir, os.path.join(FSRC, "pf_gpu.f90"),
               os.path.join(FSRC, src), "-L", HERE, "-lparafem_amd",
               "-Wl,-rpath,$ORIGIN/..", "-o", tmp]
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=moddir)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed for {src}:\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, exe)
        if verbose:
            print("built", exe)
    return P121


def build(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    build_fortran(verbose)
    shutil.rmtree(OBJDIR, ignore_errors=True)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
