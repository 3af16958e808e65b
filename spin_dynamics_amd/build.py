"""Build script: compiles the gfx950 HIP library and the amdflang driver in-tree.

    python -m spin_dynamics_amd.build            # library + Fortran driver
    python -m spin_dynamics_amd.build --lib-only

Outputs: spin_dynamics_amd/lib/libspinchem_hip.so, spin_dynamics_amd/bin/spinchem.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
LIB_DIR = os.path.join(PKG, "lib")
BIN_DIR = os.path.join(PKG, "bin")
OBJ_DIR = os.path.join(PKG, "_build")
LIB = os.path.join(LIB_DIR, "libspinchem_hip.so")
DRIVER = os.path.join(BIN_DIR, "spinchem")

ARCH = os.environ.get("SPC_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
FC = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", os.path.join(ROOT, "include"), "-Wno-unused-result"]
HIP_SOURCES = ["hamiltonian.hip", "krylov.hip", "sampling.hip", "semiclassical.hip", "adjoint.hip"]
CXX_SOURCES = ["capi.cpp", "ini.cpp", "symmetry.cpp", "newton.cpp"]
FORTRAN_SOURCES = ["spinchem_capi.f90", "spc_random.f90", "class_observables.f90", "variables.f90", "matrix_exponential.f90",
                   "dynamics.f90", "sc_dynamics.f90", "symmetry.f90", "main.f90"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_library(verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES + CXX_SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(OBJ_DIR, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") or src in ("capi.cpp", "newton.cpp") else ["-x", "c++"]
        jobs.append((src, [HIPCC, *HIP_FLAGS, *lang, "-c", path, "-o", obj], obj))
    n = min(len(jobs), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=max(n, 1)) as ex:
        for (src, cmd, _), _r in zip(jobs, ex.map(lambda j: _run(j[1]), jobs)):
            if verbose:
                print("compiled", src)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *[j[2] for j in jobs]])
    os.replace(tmp, LIB)
    return LIB


def build_driver(verbose: bool = False) -> str:
    os.makedirs(BIN_DIR, exist_ok=True)
    mod = os.path.join(OBJ_DIR, "fmod")
    os.makedirs(mod, exist_ok=True)
    objs = []
    for src in FORTRAN_SOURCES:
        path = os.path.join(FSRC, src)
        obj = os.path.join(OBJ_DIR, src + ".o")
        _run([FC, "-O2", "-module-dir", mod, "-I", mod, "-c", path, "-o", obj])
        objs.append(obj)
        if verbose:
            print("compiled", src)
    tmp = DRIVER + ".tmp"
    _run([FC, "-o", tmp, *objs, "-L", LIB_DIR, "-lspinchem_hip", f"-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(tmp, DRIVER)
    return DRIVER


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    print(build_library(a.verbose))
    if not a.lib_only and os.path.isdir(FSRC) and all(os.path.exists(os.path.join(FSRC, f)) for f in FORTRAN_SOURCES):
        print(build_driver(a.verbose))


if __name__ == "__main__":
    sys.exit(main())
