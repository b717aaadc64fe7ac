"""Builds libwgnet_hip.so in-tree with hipcc for gfx950.

    python -m wgnet_amd.build

Flags:
  -ffp-contract=off        no FMA contraction: the kernels reproduce the
                           reference's float32 operation order exactly
  -fno-gpu-flush-denormals-to-zero   IEEE denormals, as on the CPU
  -Xclang -target-feature -Xclang -packed-fp32-ops   no packed FP32 VALU ops
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libwgnet_hip.so")
SOURCES = ["wave_kernels.hip", "wave_kernels_fixed.hip", "wave_kernels_band.hip", "stream_kernels.hip",
           "misfit_kernels.hip", "precond.hip", "engine.hip"]
ARCH = os.environ.get("WGNET_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the wgnet_amd HIP library cannot be built")


def command(out: str = LIB) -> list[str]:
    return [
        hipcc(),
        f"--offload-arch={ARCH}",
        "-O3",
        "-std=c++17",
        "-ffp-contract=off",
        "-fno-gpu-flush-denormals-to-zero",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-fPIC",
        "-shared",
        "-Wall",
        *[os.path.join(CSRC, s) for s in SOURCES],
        "-o", out,
        "-Wl,-rpath,/opt/rocm/lib",
    ]


def build(verbose: bool = False) -> str:
    """Compiles every HIP source into libwgnet_hip.so (always rebuilds)."""
    tmp = LIB + ".tmp"
    cmd = command(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc failed building libwgnet_hip.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
