"""Builds libavflume.so (gfx950 kernels + C ABI + host setup) and the host driver in-tree.

    python -m avalancheflume3d_amd.build

Device code is compiled with -ffp-contract=off and IEEE division / square root so the FP64
arithmetic matches the host path bit for bit; packed FP32 instructions are disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          f"-I{INCLUDE}", f"-I{CSRC}", "-Wno-unused-command-line-argument"]

LIB = os.path.join(PKG, "libavflume.so")
DRIVER = os.path.join(PKG, "avalanche_amd")


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd[:3])} ... (exit {r.returncode})")
    noise = [ln for ln in (r.stdout + r.stderr).splitlines()
             if ln.strip() and "packed-fp32-ops" not in ln and "not a recognized feature" not in ln]
    if verbose and noise:
        print("\n".join(noise))


def build(verbose: bool = False) -> str:
    """Compiles the library and the host driver; returns the library path."""
    srcs = [os.path.join(CSRC, "dem_context.hip"), os.path.join(CSRC, "setup.cpp")]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", *COMMON, "-shared", "-o", tmp, *srcs], verbose)
    os.replace(tmp, LIB)
    drv = os.path.join(CSRC, "avalanche_main.cpp")
    if os.path.exists(drv):
        _run([HIPCC, *COMMON, "-o", DRIVER, drv, f"-L{PKG}", "-lavflume",
              f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN"], verbose)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
