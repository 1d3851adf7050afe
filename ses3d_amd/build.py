"""Build libses3d_hip.so in-tree with hipcc for gfx950.

    python -m ses3d_amd.build

Flags: -ffp-contract=off keeps every float32 multiply and add separately
rounded (the reference's operation order is reproduced bit for bit);
packed-fp32 instructions are disabled on the device.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libses3d_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SES3D_OFFLOAD_ARCH", "gfx950")

SOURCES = ["ses3d_hip.hip", "setup.cpp"]

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
          f"-I{os.path.join(ROOT, 'include')}"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {cmd[-1]}")
    return r


def build(verbose: bool = False) -> str:
    """Compile every source and link libses3d_hip.so; returns its path."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        obj = os.path.join(objdir, src + ".o")
        if src.endswith(".hip"):
            cmd = [HIPCC, "-x", "hip", *COMMON, *DEVICE, "-c", path, "-o", obj]
        else:
            cmd = [HIPCC, *COMMON, "-c", path, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", LIB + ".tmp", *objs]
    _run(cmd)
    os.replace(LIB + ".tmp", LIB)
    return LIB


FC = os.environ.get("SES3D_FC", "/opt/rocm/lib/llvm/bin/amdflang")
FORTRAN = os.path.join(HERE, "fortran")
DRIVER = os.path.join(FORTRAN, "ses3d_hip.exe")


def build_fortran(verbose: bool = False) -> str:
    """Compile the Fortran drop-in driver (iso_c_binding over libses3d_hip.so);
    returns its path, or "" when no Fortran compiler is available."""
    if not os.path.exists(FC):
        return ""
    objdir = os.path.join(HERE, "_build", "fortran")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in ("ses3d_hip_mod.f90", "ses3d_hip_main.f90"):
        obj = os.path.join(objdir, src + ".o")
        cmd = [FC, "-O2", "-J", objdir, "-c", os.path.join(FORTRAN, src), "-o", obj]
        if verbose:
            print(" ".join(cmd))
        _run(cmd)
        objs.append(obj)
    cmd = [FC, "-o", DRIVER + ".tmp", *objs, f"-L{HERE}", "-lses3d_hip", "-Wl,-rpath,$ORIGIN/.."]
    _run(cmd)
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_fortran(verbose=True))
