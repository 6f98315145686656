"""Builds libufe_dyn.so for gfx950 in-tree (ufemism_amd/_lib/).

Flags: -ffp-contract=off keeps the FP64 kernels in the reference's
operation order (no fused multiply-add), and packed FP32 instructions are
disabled for the whole library.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libufe_dyn.so")
SOURCES = ["ufe_dyn.hip", "mesh_build.hip"]
ARCH = os.environ.get("UFE_OFFLOAD_ARCH", "gfx950")


def hipcc_command(out=LIB):
    return (["hipcc", f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
             "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-o", out]
            + [os.path.join(CSRC, s) for s in SOURCES])


def build_library(verbose=False) -> str:
    os.makedirs(LIBDIR, exist_ok=True)
    tmp = LIB + ".tmp"
    cmd = hipcc_command(tmp)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose:
        msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32" not in l)
        if msg.strip():
            print(msg)
    os.replace(tmp, LIB)
    return LIB


FORTRAN_SOURCES = ["ufe_c_binding.f90", "ice_dynamics_module.f90", "ufe_step_driver.f90"]
FORTRAN_EXE = os.path.join(LIBDIR, "ufe_step_driver")


def build_fortran_host(verbose=False) -> str:
    """The Fortran host (ice_dynamics_module over ISO_C_BINDING) and its
    one-step driver, linked against libufe_dyn.so (needs flang)."""
    fdir = os.path.join(HERE, "fortran")
    mod = os.path.join(LIBDIR, "fmod")
    os.makedirs(mod, exist_ok=True)
    flang = os.environ.get("FC", "flang")
    cmd = ([flang, "-O2", "-module-dir", mod] + [os.path.join(fdir, s) for s in FORTRAN_SOURCES]
           + ["-o", FORTRAN_EXE, "-L" + LIBDIR, "-lufe_dyn", "-Wl,-rpath,$ORIGIN"])
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("flang failed:\n" + r.stdout + r.stderr)
    if verbose and (r.stdout + r.stderr).strip():
        print(r.stdout + r.stderr)
    return FORTRAN_EXE


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_fortran_host(verbose=True))
