"""Benchmark: forward + exact-adjoint shot-solves/s on one MI355X.

Workload (one step): 64 synthetic models (flat / curved / faulted families,
observed data from each model plus a Gaussian anomaly) x 5 shots on the 70 x 70
grid (114 x 114 with the 20-cell C-PML), 1000 steps of 1 ms, 15 Hz Ricker,
sources at x-cells {0, 17, 34, 52, 69}, 70 receivers.  A step is the batched
loss of the 64 models through the torch.autograd.Function (HIP forward with
snapshots, misfit, HIP adjoint) plus its backward: 320 shot-solves.

    python bench.py [--steps K] [--warmup W] [--no-cpu]

Prints one JSON line.  `roofline` is the dominant kernel's algorithmic HBM bytes
per launch over its mean launch time (HIP events on the launch stream);
`cpu_baseline` times the float32 CPU solver on the host cores on a bounded
sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

METRIC = "fwd+adjoint shot-solves/s at 70x70x1000 FP32 vs CPU ref; refine solves-to-misfit"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
CFG_SRC = np.array([[0, 10], [170, 10], [340, 10], [520, 10], [690, 10]], np.float32)


def cpu_baseline(start, true, nt, dt, threads, units):
    """The float32 CPU solver (forward + adjoint) on `units` (model, shot) solves."""
    from oracle import oracle as O

    g = O.Geometry(nt=nt, dt=dt)
    w = O.ricker_f32(nt, dt, 15.0)
    S = CFG_SRC.shape[0]
    idx = np.arange(units)
    b, s = idx // S, idx % S
    src = CFG_SRC[s]
    vel = np.ascontiguousarray(start[b].reshape(units, -1))
    obs = np.ascontiguousarray(O.batch_forward_f32(g, src, np.ascontiguousarray(true[b].reshape(units, -1)), w, threads))
    O.batch_fwd_adj_f32(g, src[:threads], vel[:threads], w, obs[:threads], threads)  # warm-up
    t0 = time.perf_counter()
    O.batch_fwd_adj_f32(g, src, vel, w, obs, threads)
    el = time.perf_counter() - t0
    return units / el, el


def traffic_from_profiles(kernel: str):
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f"traffic_{kernel}.json")
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f).get("bytes_per_launch")
    return None


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--models", type=int, default=64)
    ap.add_argument("--nt", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only; this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    import ctypes

    import torch

    from wgnet_amd import synthetic
    from wgnet_amd.physics import WavePhysics

    B, S, nt, dt = args.models, 5, args.nt, 0.001
    ph = WavePhysics(shots=S, nt=nt, dt=dt)
    ph.engine.set_sources(CFG_SRC)
    start, true = synthetic.batch(B, seed=1000)
    start_km = torch.from_numpy(start / 1000.0).cuda()
    obs = ph.simulate_torch(torch.from_numpy(true / 1000.0).cuda())
    fields = start_km.clone().requires_grad_(True)

    def step():
        losses = ph.misfit(fields, obs)
        losses.sum().backward()
        fields.grad = None
        return losses

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    lib, h = ph.engine._lib, ph.engine.handle
    fms, fn, ams, an = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double(), ctypes.c_int64()
    lib.wg_timing_enable(h, 1)
    lib.wg_timing_read(h, ctypes.byref(fms), ctypes.byref(fn), ctypes.byref(ams), ctypes.byref(an))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        losses = step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    lib.wg_timing_read(h, ctypes.byref(fms), ctypes.byref(fn), ctypes.byref(ams), ctypes.byref(an))
    lib.wg_timing_enable(h, 0)
    assert bool(torch.isfinite(losses).all())

    units = B * S
    value = units * args.steps / el
    nx, nz, _ = ph.engine.dims()
    inner, R = (nx - 4) * (nz - 4), ph.engine.nrec
    # Algorithmic HBM bytes per unit: forward writes the snapshot (4 B per inner
    # cell-step) and the gathers (4 B per receiver-step); the adjoint reads the
    # snapshot and the residual.
    fwd_bytes = (4 * inner * nt + 4 * R * nt) * units
    adj_bytes = (4 * inner * nt + 4 * R * nt) * units
    f_avg = fms.value / max(fn.value, 1) / 1e3
    a_avg = ams.value / max(an.value, 1) / 1e3
    if a_avg >= f_avg:
        kern, kbytes, kavg = "k_adjoint", adj_bytes, a_avg
    else:
        kern, kbytes, kavg = "k_forward", fwd_bytes, f_avg
    achieved = kbytes / kavg / 1e9
    # Per shot-solve algorithmic figures (bytes 8 inner nt + 12 R nt).
    ss_bytes = 8 * inner * nt + 12 * R * nt
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "shot-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * el / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (flat/curved/faulted 3-layer models, Gaussian-anomaly observed data)",
        "config": {
            "workload": f"{B} models x {S} shots, 70x70 grid (114x114 padded), nt={nt}, dt=1 ms, "
                        "forward + adjoint gradient through torch.autograd.Function",
            "models": B, "shots": S, "nt": nt, "grid": "70x70", "pml": 20, "parallelism": "1 GPU",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kern,
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(kern),
            "algorithmic_bytes_per_launch": kbytes,
            "launch_ms": {"k_forward": round(1e3 * f_avg, 3), "k_adjoint": round(1e3 * a_avg, 3)},
            "shot_solve_hbm_frac": round(ss_bytes * value / 1e9 / HBM_PEAK_GBS, 4),
        },
    }
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        units_cpu = 2 * threads
        rate, cel = cpu_baseline(start, true, nt, dt, threads, units_cpu)
        out["cpu_baseline"] = {
            "value": round(rate, 3),
            "unit": "shot-solves/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{units_cpu} (model, shot) forward+adjoint solves of the same workload, "
                      f"{threads} threads, {cel:.1f} s wall; float32 C restatement of the reference solver",
        }
        out["speedup_vs_cpu"] = round(value / rate, 1)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
