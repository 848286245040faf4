"""Builds libsph2d.so (gfx950) in-tree with hipcc.

Translation units: the FP64 step and the FP64 force evaluation (no FMA
contraction, for parity with the reference build), the FP32-mixed force
evaluation (contraction within each expression only, so every kernel that
inlines the same pair terms rounds them alike), the C-ABI/host driver and the host-side
writers of the reference's file formats. Packed FP32
instructions are disabled on every HIP line.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libsph2d.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SPH2D_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-command-line-argument",
]

UNITS = [
    ("sph2d_fp64.hip", ["-ffp-contract=off"]),
    ("sph2d_eval64.hip", ["-ffp-contract=off"]),
    ("sph2d_eval32.hip", ["-ffp-contract=on"]),
    ("sph2d_reorder.hip", ["-ffp-contract=off"]),
    ("sph2d_capi.hip", ["-ffp-contract=off"]),
    ("sph2d_output.hip", ["-ffp-contract=off"]),
    ("sph2d_adjoint.hip", ["-ffp-contract=off"]),
    ("sph2d_adjstep.hip", ["-ffp-contract=off"]),
    ("sph2d_ms.hip", ["-ffp-contract=off"]),
]

HEADERS = ["sph2d_internal.h", "sph2d_device.h", "sph2d_riemann.h"]


def _run(cmd):
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + p.stdout + p.stderr)
    return p


FC = os.environ.get("FC", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN = os.path.join(HERE, "fortran")


def build_fortran(verbose: bool = False) -> str:
    """The ISO_C_BINDING interface module and the Fortran drivers (the whole step,
    and the step routine by routine), linked against libsph2d.so (skipped when
    flang is absent)."""
    if not os.path.exists(FC):
        return ""
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    for prog in ("tgv_driver", "routine_driver"):
        exe = os.path.join(FORTRAN, prog)
        cmd = [FC, "-O2", "-J", bdir, "-o", exe,
               os.path.join(FORTRAN, "sph2d_interface.f90"), os.path.join(FORTRAN, prog + ".f90"),
               "-L" + HERE, "-lsph2d", "-Wl,-rpath,$ORIGIN/.."]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
    return exe


def build(verbose: bool = False) -> str:
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    for src, extra in UNITS:
        obj = os.path.join(bdir, src.replace(".hip", ".o"))
        cmd = [HIPCC, *COMMON, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    _run(cmd)
    os.replace(tmp, OUT)
    build_fortran(verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
