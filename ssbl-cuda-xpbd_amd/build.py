"""Compiles libssbl_xpbd_hip_abi41.so for gfx950 with hipcc.

All device code is built with -ffp-contract=off (no implicit multiply-add
fusion: every fused multiply-add in the kernels is an explicit fmaf), IEEE
single-precision division and square root (hipcc's default), denormals kept,
and packed FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SOURCES = ["csrc/xpbd_kernels.hip", "csrc/xpbd_self.hip", "csrc/xpbd_pcg.hip", "csrc/xpbd_dynamic.hip", "csrc/xpbd_solver.hip"]
OUTPUT = PKG / "libssbl_xpbd_hip_abi41.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-value",
]


def build(verbose: bool = False) -> Path:
    srcs = [str(PKG / s) for s in SOURCES if (PKG / s).exists()]
    tmp = OUTPUT.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, f"-I{ROOT / 'include'}", f"-I{PKG / 'csrc'}", "-o", str(tmp), *srcs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stderr[-8000:]}")
    if verbose and res.stderr:
        print(res.stderr)
    os.replace(tmp, OUTPUT)
    return OUTPUT


if __name__ == "__main__":
    print(build(verbose=True))
