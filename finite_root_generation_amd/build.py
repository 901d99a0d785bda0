"""Builds libfrg.so (the HIP tracker + C ABI) in-tree for gfx950.

    python -m finite_root_generation_amd.build        # or: python finite_root_generation_amd/build.py

hipcc compiles the device code for gfx950 without a GPU present.  Packed FP32 instructions are
disabled for all device code (``-target-feature -packed-fp32-ops``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libfrg.so")
ARCH = os.environ.get("FRG_OFFLOAD_ARCH", "gfx950")

SOURCES = ["frg_tracker.hip", "frg_mlp.hip", "frg_dedupe.hip", "frg_capi.cpp", "frg_host.cpp", "frg_flops.cpp"]
HEADERS = ["frg_device.h", "frg_system.h", "frg_kernels.h"]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def compile_cmd(out: str = LIB) -> list:
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument", "-I", os.path.join(ROOT, "include")]
    cmd += [os.path.join(CSRC, s) for s in SOURCES]
    cmd += ["-o", out]
    return cmd


ROUNDS_SRC = os.path.join(PKG, "host", "frg_rounds.cpp")
ROUNDS_EXE = os.path.join(PKG, "frg_rounds")


def rounds_cmd(out: str = ROUNDS_EXE) -> list:
    """The round driver: plain C++ over the C ABI, linked against the in-tree libfrg.so."""
    rocm_lib = os.path.join(os.path.dirname(os.path.dirname(hipcc())), "lib")
    return [os.environ.get("CXX", "g++"), "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
            ROUNDS_SRC, "-L", PKG, "-lfrg", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath-link,{rocm_lib}", "-o", out]


def build(verbose: bool = False) -> str:
    tmp = LIB + ".tmp"
    cmd = compile_cmd(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    # hipcc's host pass does not know the packed-fp32 feature and says so; that is harmless
    noise = "is not a recognized feature for this target"
    msgs = [ln for ln in (r.stderr or "").splitlines() if noise not in ln]
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + "\n".join(msgs))
    if verbose and msgs:
        print("\n".join(msgs))
    os.replace(tmp, LIB)
    r = subprocess.run(rounds_cmd(ROUNDS_EXE + ".tmp"), capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("building frg_rounds failed:\n" + r.stderr)
    os.replace(ROUNDS_EXE + ".tmp", ROUNDS_EXE)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
