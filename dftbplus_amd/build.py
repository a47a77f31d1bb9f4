"""Compile the HIP sources into the in-tree shared library ``libdftbhip.so`` (gfx950).

Each translation unit is compiled in parallel with hipcc and then linked; the packed-FP32
target feature is disabled for every device compilation.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libdftbhip.so")
SOURCES = ["gemm.hip", "dense.hip", "sytrd.hip", "band.hip", "stedc.hip", "eigen.hip", "sparse.hip",
           "batch.hip", "batch_scc.hip", "neigh.hip", "zhegv.hip", "capi.hip"]
ARCH = os.environ.get("DFTBHIP_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I", os.path.join(ROOT, "include")]


def _compile(src: str, objdir: str) -> str:
    obj = os.path.join(objdir, src.replace(".hip", ".o"))
    cmd = ["hipcc", *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> str:
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(s, objdir), SOURCES))
    tmp = LIB + ".tmp"
    cmd = ["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}", file=sys.stderr)
    build_fortran(objdir, verbose)
    return LIB


FORTRAN_TEST = os.path.join(HERE, "fortran", "test_dftbhip")


def build_fortran(objdir: str, verbose: bool = True) -> None:
    """Compile the ISO_C_BINDING module and its test program with amdflang when present."""
    import shutil
    fc = shutil.which("amdflang") or (os.path.exists("/opt/rocm/bin/amdflang") and "/opt/rocm/bin/amdflang")
    if not fc:
        return
    fdir = os.path.join(HERE, "fortran")
    mod_o = os.path.join(objdir, "dftbhip_mod.o")
    r = subprocess.run([fc, "-O2", "-c", os.path.join(fdir, "dftbhip.F90"), "-o", mod_o, "-J", objdir],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed for dftbhip.F90:\n{r.stderr}")
    r = subprocess.run([fc, "-O2", os.path.join(fdir, "test_dftbhip.F90"), mod_o, "-I", objdir,
                        "-L" + HERE, "-ldftbhip", "-Wl,-rpath,$ORIGIN/..", "-o", FORTRAN_TEST],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"amdflang failed for test_dftbhip.F90:\n{r.stderr}")
    if verbose:
        print(f"built {FORTRAN_TEST}", file=sys.stderr)


if __name__ == "__main__":
    build()
