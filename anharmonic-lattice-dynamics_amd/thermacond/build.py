"""Builds libalt3.so (HIP kernels + C ABI) for gfx950 with hipcc, in-tree, and the Fortran
job-1 driver fortran/alt3_kappa over it with amdflang.

    python -m thermacond.build            (with anharmonic-lattice-dynamics_amd/ on sys.path)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(PKG_DIR, "csrc")
INCLUDE = os.path.join(os.path.dirname(PKG_DIR), "include")
LIB_PATH = os.path.join(PKG_DIR, "libalt3.so")
OBJ_DIR = os.path.join(PKG_DIR, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["host_tables.cpp", "isotopes.cpp", "capi.cpp", "eig.hip", "v3.hip", "v3_mixed.hip", "v3_rows.hip", "v3_chain.hip", "v3_chain16.hip", "fgr.hip", "bte.hip",
           "self.hip", "cg_static.hip"]
FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I" + INCLUDE, "-I" + CSRC]


def _compile(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [HIPCC] + FLAGS + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


FORTRAN_DIR = os.path.join(PKG_DIR, "fortran")
FORTRAN_EXE = os.path.join(FORTRAN_DIR, "alt3_kappa")
FLANG = os.environ.get("FLANG", "/opt/rocm/bin/amdflang")


def build_fortran(verbose: bool = True) -> str:
    """The Fortran job-1 driver (fortran/alt3_kappa.f90 over the iso_c_binding module
    fortran/alt3_iface.f90), linked against libalt3.so with an rpath relative to itself."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    objs = []
    for src in ("alt3_iface.f90", "alt3_kappa.f90"):
        obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".f.o")
        cmd = [FLANG, "-O2", "-J", OBJ_DIR, "-c", os.path.join(FORTRAN_DIR, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed for {src}:\n{r.stdout}\n{r.stderr}")
        objs.append(obj)
    tmp = FORTRAN_EXE + ".tmp"
    cmd = [FLANG, "-o", tmp] + objs + ["-L" + PKG_DIR, "-lalt3", "-Wl,-rpath,$ORIGIN/.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link of alt3_kappa failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, FORTRAN_EXE)
    if verbose:
        print(f"built {FORTRAN_EXE}")
    return FORTRAN_EXE


def build(verbose: bool = True) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"built {LIB_PATH}")
    build_fortran(verbose)
    return LIB_PATH


if __name__ == "__main__":
    build()
    sys.exit(0)
