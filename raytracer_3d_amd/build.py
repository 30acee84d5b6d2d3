"""Build the native pieces in-tree.

* ``librt6d.so`` -- the HIP library (raytracer_3d_amd/csrc/*) for gfx950,
  exporting the C ABI of include/rt6d.h.
* ``fortran/`` -- the Fortran drop-in driver linked against librt6d.so
  (see raytracer_3d_amd/fortran/Makefile), with the package's case-file
  plug-in or with a config plug-in of the reference's form.

Run ``python -m raytracer_3d_amd.build`` or call :func:`build_all`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "librt6d.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("RT6D_ARCH", "gfx950")

# -packed-fp32-ops: never emit packed binary32 add/mul/fma; binary32 code is
# scalar on the VALU.  -ffp-contract=off in the exact translation unit keeps
# the reference's separate rounding of every operation.
COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]

# -fno-honor-nans in the fast unit: the states are finite, so min/max need
# not quiet signalling NaNs first (one instruction less per face limiter).
UNITS = [
    ("rt6d_exact.hip", ["-ffp-contract=off"]),
    ("rt6d_fast.hip", ["-ffp-contract=fast", "-fno-honor-nans"]),
    ("rt6d_api.cpp", ["-x", "hip"]),
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("command failed: " + " ".join(cmd))
    return r


def build_library(verbose: bool = False) -> str:
    bdir = os.path.join(PKG, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    for src, extra in UNITS:
        obj = os.path.join(bdir, src.rsplit(".", 1)[0] + ".o")
        cmd = [HIPCC] + COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose: bool = False) -> str | None:
    """Build the Fortran driver (needs amdflang); returns the binary path."""
    fdir = os.path.join(PKG, "fortran")
    if not os.path.exists(os.path.join(fdir, "Makefile")):
        return None
    if shutil.which("amdflang") is None and not os.path.exists("/opt/rocm/bin/amdflang"):
        return None
    r = subprocess.run(["make", "-s", "-C", fdir], capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("fortran build failed")
    return os.path.join(fdir, "rt6d_driver")


def build_plugin(src: str, outdir: str, mpi: bool = False) -> str | None:
    """Build the Fortran driver around a config plug-in (the reference's
    for_config/*.f90 form) as outdir/model_<name>.x; None without amdflang."""
    fdir = os.path.join(PKG, "fortran")
    if shutil.which("amdflang") is None and not os.path.exists("/opt/rocm/bin/amdflang"):
        return None
    cmd = ["make", "-s", "-C", fdir, "plugin", f"PLUGIN={os.path.abspath(src)}",
           f"PLUGIN_DIR={os.path.abspath(outdir)}", f"MPI={1 if mpi else 0}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("plug-in build failed: " + src)
    return os.path.join(outdir, "model_" + os.path.splitext(os.path.basename(src))[0] + ".x")


def build_all(verbose: bool = False) -> None:
    build_library(verbose)
    build_fortran(verbose)


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", LIB)
