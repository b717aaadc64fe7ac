"""Builds libwgnet_hip.so in-tree with hipcc for gfx950.

    python -m wgnet_amd.build

Each source compiles to its own object in parallel (wgnet_amd/_build/), then
hipcc links the shared library.

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
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libwgnet_hip.so")
SOURCES = ["wave_kernels.hip", "wave_kernels_fixed.hip", "wave_kernels_fixed_fwd.hip", "wave_kernels_band.hip",
           "wave_kernels_fused.hip", "stream_kernels.hip", "misfit_kernels.hip", "precond.hip", "engine.hip"]
# Per-source flags: the resident 70 x 70 forward bodies run faster with the
# machine scheduler's occupancy-first strategy (the adjoint's spill with it),
# the one-launch uneven batch with the max-ilp strategy (the band kernels of
# small batches are slower with it).
SOURCE_FLAGS = {"wave_kernels_fixed_fwd.hip": ["-mllvm", "--amdgpu-sched-strategy=iterative-maxocc"],
                "wave_kernels_fused.hip": ["-mllvm", "--amdgpu-sched-strategy=max-ilp"]}
ARCH = os.environ.get("WGNET_OFFLOAD_ARCH", "gfx950")
FLAGS = [
    "-O3",
    "-std=c++17",
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fPIC",
    "-Wall",
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the wgnet_amd HIP library cannot be built")


def compile_command(src: str) -> list[str]:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    return [hipcc(), f"--offload-arch={ARCH}", *FLAGS, *SOURCE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src),
            "-o", obj]


def link_command(out: str = LIB) -> list[str]:
    objs = [os.path.join(OBJ, os.path.splitext(s)[0] + ".o") for s in SOURCES]
    return [hipcc(), f"--offload-arch={ARCH}", "-fPIC", "-shared", *objs, "-o", out, "-Wl,-rpath,/opt/rocm/lib"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed: {' '.join(cmd[-3:])}")


def build(verbose: bool = False) -> str:
    """Compiles every HIP source and links libwgnet_hip.so (always rebuilds)."""
    os.makedirs(OBJ, exist_ok=True)
    jobs = max(1, min(len(SOURCES), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(jobs) as pool:
        for f in [pool.submit(_run, compile_command(s), verbose) for s in SOURCES]:
            f.result()
    tmp = LIB + ".tmp"
    _run(link_command(tmp), verbose)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
