"""Benchmark: forward + exact-adjoint shot-solves/s on one MI355X.

Workload (one step): 64 synthetic models (flat / curved / faulted families,
observed data from each model plus a Gaussian anomaly) x 5 shots on the 70 x 70
grid (114 x 114 with the 20-cell C-PML), 1000 steps of 1 ms, 15 Hz Ricker,
sources at x-cells {0, 17, 34, 52, 69}, 70 receivers.  A step is the batched
loss of the 64 models through the torch.autograd.Function (HIP forward with
snapshots, misfit, HIP adjoint) plus its backward: 320 shot-solves.

    python bench.py [--steps K] [--warmup W] [--no-cpu]
    python bench.py --workload large [--steps K] [--warmup W] [--no-cpu]

`--workload large` measures the HBM-streamed path instead: 8 models of
1000 x 300 cells (1044 x 344 padded), 20 shots, 1000 receivers, 4000 steps of
0.8 ms, adjoint with a checkpoint every 100 steps and segment recompute; it
also reports the FP64 dot-product test of the wavelet operator at that size.

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
SURFACE_SRC = np.array([[0, 10], [170, 10], [340, 10], [520, 10], [690, 10]], np.float32)


def cpu_baseline(start, true, nt, dt, threads, units):
    """The float32 CPU solver (forward + adjoint) on `units` (model, shot) solves."""
    from oracle import oracle as O

    g = O.Geometry(nt=nt, dt=dt)
    w = O.ricker_f32(nt, dt, 15.0)
    S = SURFACE_SRC.shape[0]
    idx = np.arange(units)
    b, s = idx // S, idx % S
    src = SURFACE_SRC[s]
    vel = np.ascontiguousarray(start[b].reshape(units, -1))
    obs = np.ascontiguousarray(O.batch_forward_f32(g, src, np.ascontiguousarray(true[b].reshape(units, -1)), w, threads))
    O.batch_fwd_adj_f32(g, src[:threads], vel[:threads], w, obs[:threads], threads)  # warm-up
    t0 = time.perf_counter()
    O.batch_fwd_adj_f32(g, src, vel, w, obs, threads)
    el = time.perf_counter() - t0
    return units / el, el


def cpu_baseline_large(start, src, nvx, nvz, nt, dt, threads, units):
    """The float32 CPU solver (forward + adjoint) on `units` large-grid solves."""
    from oracle import oracle as O

    g = O.Geometry(nvx=nvx, nvz=nvz, nt=nt, dt=dt)
    w = O.ricker_f32(nt, dt, 15.0)
    idx = np.arange(units)
    s = src[idx % src.shape[0]]
    vel = np.ascontiguousarray(np.repeat(start[:1].reshape(1, -1), units, 0))
    obs = np.ascontiguousarray(O.batch_forward_f32(g, s, vel * np.float32(1.02), w, threads))
    t0 = time.perf_counter()
    O.batch_fwd_adj_f32(g, s, vel, w, obs, threads)
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
    ap.add_argument("--workload", choices=("batch70", "large", "copy"), default="batch70")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only; this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2
    if args.workload == "large":
        return bench_large(args)
    if args.workload == "copy":
        return bench_copy(args)

    import ctypes

    import torch

    from wgnet_amd import synthetic
    from wgnet_amd.physics import WavePhysics

    B, S, nt, dt = args.models, 5, args.nt, 0.001
    ph = WavePhysics(shots=S, nt=nt, dt=dt)
    ph.engine.set_sources(SURFACE_SRC)
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
        kern, kbytes, kavg = "k_adjoint_fx", adj_bytes, a_avg
    else:
        kern, kbytes, kavg = "k_forward_fx", fwd_bytes, f_avg
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
            "launch_ms": {"k_forward_fx": round(1e3 * f_avg, 3), "k_adjoint_fx": round(1e3 * a_avg, 3)},
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


def bench_copy(args) -> int:
    """Calibration for hardware traffic counters: `steps` device-to-device
    copies of 1 GiB (torch's copy kernel reads and writes 1 GiB each)."""
    import torch

    n = 1 << 28
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    gbs = 2 * 4 * n * args.steps / el / 1e9
    print(json.dumps({"workload": "copy", "bytes_read_per_launch": 4 * n, "bytes_written_per_launch": 4 * n,
                      "launches": args.steps + args.warmup, "GB/s": round(gbs, 1)}), flush=True)
    return 0


def bench_large(args) -> int:
    """Streamed 1000 x 300 x 4000 workload (checkpoint-every-100 adjoint)."""
    import ctypes

    import torch

    from wgnet_amd import _lib, synthetic
    from wgnet_amd.physics import WavePhysics, _dev_ptr, _stream

    B, S, nvx, nvz, nt, dt = 8, 20, 1000, 300, args.nt if args.nt != 1000 else 4000, 0.0008
    ph = WavePhysics(shots=S, nt=nt, dt=dt, nvx=nvx, nvz=nvz)
    eng = ph.engine
    assert eng.streamed
    src = np.stack([np.linspace(0.0, (nvx - 1) * 10.0, S), np.full(S, 10.0)], 1).astype(np.float32)
    eng.set_sources(src)
    start, true = synthetic.large_batch(B, seed=3000, nvz=nvz, nvx=nvx)
    obs = ph.simulate_torch(torch.from_numpy(true / 1000.0).cuda())
    fields = torch.from_numpy(start / 1000.0).cuda().requires_grad_(True)

    def step():
        losses = ph.misfit(fields, obs)
        losses.sum().backward()
        fields.grad = None
        return losses

    for _ in range(args.warmup):
        print(f"large: warmup step", file=sys.stderr, flush=True)
        step()
    torch.cuda.synchronize()
    lib, h = eng._lib, eng.handle
    fms, fn, ams, an = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double(), ctypes.c_int64()
    lib.wg_timing_enable(h, 1)
    lib.wg_timing_read(h, ctypes.byref(fms), ctypes.byref(fn), ctypes.byref(ams), ctypes.byref(an))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        losses = step()
        torch.cuda.synchronize()
        print(f"large: step {k} done", file=sys.stderr, flush=True)
    el = time.perf_counter() - t0
    lib.wg_timing_read(h, ctypes.byref(fms), ctypes.byref(fn), ctypes.byref(ams), ctypes.byref(an))
    lib.wg_timing_enable(h, 0)
    assert bool(torch.isfinite(losses).all())

    # FP64 dot-product test of the wavelet operator F (source forcing -> traces)
    # on one model: <F x, F x> = <x, F^T F x>.
    rng = np.random.default_rng(7)
    x = torch.from_numpy(rng.standard_normal((1, S, nt)).astype(np.float32)).cuda()
    v1 = torch.from_numpy(start[:1].copy()).cuda()
    fx = torch.empty((1, S, nt, eng.nrec), device="cuda")
    _lib.check(lib.wg_simulate_wavelets_dev(h, 1, _dev_ptr(v1), _dev_ptr(x), _dev_ptr(fx), _stream()))
    ftfx = torch.empty((1, S, nt), device="cuda")
    _lib.check(lib.wg_adjoint_apply_dev(h, 1, _dev_ptr(v1), _dev_ptr(fx), None, _dev_ptr(ftfx), None, _stream()))
    lhs, rhs = float((fx.double() ** 2).sum()), float((x.double() * ftfx.double()).sum())
    dot_rel = abs(lhs - rhs) / abs(lhs)
    del fx

    units = B * S
    value = units * args.steps / el
    nx, nz, _ = eng.dims()
    inner = (nx - 4) * (nz - 4)
    f_avg = fms.value / max(fn.value, 1) / 1e3
    a_avg = ams.value / max(an.value, 1) / 1e3
    # Algorithmic HBM bytes per inner cell-step of one unit: adjoint 20 (lambda
    # t+1 and t read, lambda t-1 written, v and snapshot read); forward 16
    # (u t-1 and t read, u t+1 written, v read), 20 when it stores the snapshot.
    adj_bytes = 20 * inner * units
    achieved = adj_bytes / a_avg / 1e9
    ss_bytes = 56 * inner * nt  # forward + recompute + adjoint per shot-solve
    out = {
        "metric": "fwd+adjoint shot-solves/s, HBM-streamed 1000x300x4000 FP32, checkpoint every 100 steps",
        "value": round(value, 4),
        "unit": "shot-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * el / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (8 curved layers + 20 Gaussian anomalies per model)",
        "config": {
            "workload": f"{B} models x {S} shots, {nvx}x{nvz} grid ({nx}x{nz} padded), nt={nt}, dt=0.8 ms, "
                        "1000 receivers, forward + checkpointed adjoint gradient through torch.autograd.Function",
            "models": B, "shots": S, "nt": nt, "grid": f"{nvx}x{nvz}", "pml": 20, "checkpoint_every": 100,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "k_adj_step",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles("k_adj_step"),
            "algorithmic_bytes_per_launch": adj_bytes,
            "launch_ms": {"k_fwd_step": round(1e3 * f_avg, 4), "k_adj_step": round(1e3 * a_avg, 4)},
            "shot_solve_hbm_frac": round(ss_bytes * value / 1e9 / HBM_PEAK_GBS, 4),
        },
        "dot_test": {"operator": "wavelet -> traces, 1 model x 20 shots, full nt", "rel": dot_rel},
    }
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        nt_cpu = 500
        units_cpu = 2 * threads
        rate, cel = cpu_baseline_large(start, src, nvx, nvz, nt_cpu, dt, threads, units_cpu)
        out["cpu_baseline"] = {
            "value": round(rate * nt_cpu / nt, 5),
            "unit": "shot-solves/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{units_cpu} (model, shot) forward+adjoint solves on the 1000x300 grid at nt={nt_cpu}, "
                      f"{threads} threads, {cel:.1f} s wall, rate scaled by {nt_cpu}/{nt} (work is linear in nt); "
                      "float32 C restatement of the reference solver",
        }
        out["speedup_vs_cpu"] = round(value / out["cpu_baseline"]["value"], 1)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
