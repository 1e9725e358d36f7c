"""Builds libnssdmrg.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m nssdmrg_amd.build_lib   (or nssdmrg_amd.build_lib.build()).
Objects go to nssdmrg_amd/_build/, the library to nssdmrg_amd/libnssdmrg.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnssdmrg.so")
SOURCES = ["nss_kernels.hip", "nss_rho_big.hip", "nss_superblock.cpp", "nss_lanczos.cpp", "nss_eig.cpp", "nss_engine.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed FP32 (v_pk_*_f32) generation is disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                # MFMA accumulators in VGPRs: no AGPR round trips of loop-carried tiles
                "-mllvm", "-amdgpu-mfma-vgpr-form"]
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-result",
          "-Wno-unused-command-line-argument"]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *COMMON, *DEVICE_FLAGS, "-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    errs = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                     if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{errs}")
    if errs:
        print(errs, file=sys.stderr)
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), 8)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
           "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/llvm/bin/flang")
FORTRAN_SRC = os.path.join(HERE, "fortran", "nss_dmrg.f90")
FORTRAN_OBJ = os.path.join(HERE, "fortran", "_build")


def build_fortran(verbose: bool = False) -> str | None:
    """Compiles the Fortran binding module (nss_dmrg) and the Fortran test
    driver tests/fortran/spin_chain_golden.f90 with flang, linked against
    libnssdmrg.so through an $ORIGIN-relative rpath.  Returns the driver
    path, or None when flang is not installed."""
    if not os.path.exists(FLANG):
        return None
    root = os.path.dirname(HERE)
    tdir = os.path.join(root, "tests", "fortran")
    tbuild = os.path.join(tdir, "_build")
    os.makedirs(FORTRAN_OBJ, exist_ok=True)
    os.makedirs(tbuild, exist_ok=True)
    obj = os.path.join(FORTRAN_OBJ, "nss_dmrg.o")
    r = subprocess.run([FLANG, "-O2", "-c", f"-J{FORTRAN_OBJ}", FORTRAN_SRC, "-o", obj], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"flang failed on {FORTRAN_SRC}:\n{r.stdout}{r.stderr}")
    exe = os.path.join(tbuild, "spin_chain_golden")
    rpath = "-Wl,-rpath,$ORIGIN/" + os.path.relpath(HERE, tbuild)
    r = subprocess.run([FLANG, "-O2", f"-I{FORTRAN_OBJ}", os.path.join(tdir, "spin_chain_golden.f90"), obj,
                        "-o", exe, f"-L{HERE}", "-lnssdmrg", rpath], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"flang failed on the Fortran driver:\n{r.stdout}{r.stderr}")
    if verbose:
        print(f"built {exe}")
    return exe


if __name__ == "__main__":
    build(verbose=True)
    build_fortran(verbose=True)
