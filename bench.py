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

`--workload uneven` runs 512 uneven solves as one HIP graph, `--workload refine`
the learned preconditioner inside projected L-BFGS, and `--workload wavecache`
the training-data cache writer (2560 models, nt = 2001, decimated float16 file).

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
VALU_CLOCK_HZ = 2.4e9  # MI355X peak engine clock
FP32_PEAK_TFLOPS = 157.3
SURFACE_SRC = np.array([[0, 10], [170, 10], [340, 10], [520, 10], [690, 10]], np.float32)


def host_cpu() -> dict:
    """The host's CPU model and core counts (all cores, and those this process
    may run on)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "host_cores": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0))}


def relerr(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def cpu_baseline(start, true, nt, dt, threads, units, check=None):
    """The float32 CPU solver (forward + adjoint) on `units` (model, shot) solves.

    With `check` (the timed batch's last losses, gradients and gathers), the
    CPU leg also solves `check["models"]` with the oracle in float32 and
    float64 and returns the GPU's errors against both."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle import oracle as O

    g = O.Geometry(nt=nt, dt=dt)
    w = O.ricker_f32(nt, dt, 15.0)
    S = SURFACE_SRC.shape[0]
    idx = np.arange(units)
    b, s = (idx // S) % start.shape[0], idx % S
    src = SURFACE_SRC[s]
    vel = np.ascontiguousarray(start[b].reshape(units, -1))
    obs = np.ascontiguousarray(O.batch_forward_f32(g, src, np.ascontiguousarray(true[b].reshape(units, -1)), w, threads))
    O.batch_fwd_adj_f32(g, src[:threads], vel[:threads], w, obs[:threads], threads)  # warm-up
    t0 = time.perf_counter()
    O.batch_fwd_adj_f32(g, src, vel, w, obs, threads)
    el = time.perf_counter() - t0
    if check is None:
        return units / el, el, None

    wv = check["wavelet"]

    def one(m):
        v, tv, ob = check["start"][m], check["true"][m], check["obs"][m]
        l32, g32 = O.evaluate_adjoint_f32(g, SURFACE_SRC, v, wv, ob)
        tr32, tr64, ob64 = [], [], []
        for x, z in SURFACE_SRC:
            gs = O.Geometry(nt=nt, dt=dt, src_x=float(x), src_z=float(z))
            tr32.append(O.forward_f32(gs, v, wv))
            tr64.append(O.forward_f64(gs, v, wv.astype(np.float64)))
            ob64.append(O.forward_f64(gs, tv, wv.astype(np.float64)))
        ob64 = np.stack(ob64)
        l64, g64 = O.evaluate_adjoint_f64(g, SURFACE_SRC, v, wv.astype(np.float64), ob64)
        return m, l32, g32, l64, g64, np.stack(tr32), np.stack(tr64), ob64

    with ThreadPoolExecutor(min(threads, len(check["models"]))) as pool:
        res = list(pool.map(one, check["models"]))
    worst = {"gathers_f32": 0.0, "loss_f32": 0.0, "grad_f32": 0.0}
    per_model = []
    bit_equal = True
    for m, l32, g32, l64, g64, tr32, tr64, ob64 in res:
        loss, gk, tr = float(check["losses"][m]), check["grads_km"][m], check["traces"][m]
        gms = gk.astype(np.float64) / 1000.0
        worst["gathers_f32"] = max(worst["gathers_f32"], relerr(tr, tr32))
        worst["loss_f32"] = max(worst["loss_f32"], abs(loss - l32) / max(l32, 1e-300))
        # in the device's units (km/s: the float32 oracle gradient times 1000, as the engine scales it)
        worst["grad_f32"] = max(worst["grad_f32"], relerr(gk, g32 * np.float32(1000.0)))
        eps = max(relerr(tr, tr64), relerr(check["obs"][m], ob64))
        per_model.append({"model": int(m), "family": ("flat", "curved", "faulted")[m % 3], "misfit": loss,
                          "gathers_rel_f64": float(f"{relerr(tr, tr64):.3e}"),
                          "loss_rel_f64": float(f"{abs(loss - l64) / max(l64, 1e-300):.3e}"),
                          "grad_rel_f64": float(f"{relerr(gms, g64):.3e}"),
                          "kappa": float(f"{eps / np.sqrt(max(2.0 * loss, 1e-300)):.3e}")})
        bit_equal &= bool(np.array_equal(tr, tr32)) and loss == l32 and bool(
            np.array_equal(gk, g32 * np.float32(1000.0)))
    chk = {"what": "the timed batch's last losses and gradients, and the gathers of the same call, against the "
                   "CPU oracle on the same inputs (float32) and a float64 run of the same problem; models: the "
                   "largest misfit of each family and the last model",
           "max_rel_err_f32": {k: float(f"{v:.3e}") for k, v in worst.items()},
           "bit_equal_f32": bit_equal, "repeat_bit_equal": check["repeat_equal"],
           "gates_f32": {"gathers": 1e-5, "loss": 1e-6, "grad": 1e-6},
           "f64": per_model,
           "kappa": "float32 conditioning of the residual: gathers' float32 error / sqrt(2 misfit)"}
    chk["pass"] = bool(worst["gathers_f32"] <= 1e-5 and worst["loss_f32"] <= 1e-6 and worst["grad_f32"] <= 1e-6
                       and check["repeat_equal"])
    return units / el, el, chk


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


def profile_figure(kind: str, kernel: str, key: str, run_cfg: dict):
    """(value, source) of `key` in profiles/<kind>_<kernel>.json when that
    profile was taken on this run's workload (its "config" equals run_cfg),
    else None: a counter read on another configuration is not this run's."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f"{kind}_{kernel}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if d.get("config") != run_cfg or key not in d:
        return None
    return d[key], f"profiles/{kind}_{kernel}.json (rocprofv3 --pmc, same workload, measured before this run)"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=150)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--models", type=int, default=64)
    ap.add_argument("--nt", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-refine", action="store_true", help="default workload: skip the in-run refine "
                    "solves-to-misfit measurement (about two minutes)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--workload", choices=("batch70", "large", "copy", "refine", "uneven", "wavecache", "physstage",
                                           "fwi"),
                    default="batch70")
    ap.add_argument("--cache-models", type=int, default=2560)
    ap.add_argument("--cache-batch", type=int, default=256)
    ap.add_argument("--cache-loop-models", type=int, default=64)
    ap.add_argument("--phys-batch", type=int, default=8)
    ap.add_argument("--ref-steps", type=int, default=2)
    ap.add_argument("--fwi-models", type=int, default=24)
    ap.add_argument("--fwi-budget", type=int, default=120)
    ap.add_argument("--solves", type=int, default=512)
    ap.add_argument("--train-pairs", type=int, default=20000)
    ap.add_argument("--eval-models", type=int, default=256)
    ap.add_argument("--budget", type=int, default=160)
    ap.add_argument("--epochs", type=int, default=100)
    ap.add_argument("--save-params", default="")
    ap.add_argument("--pc-mode", choices=("jump", "first", "replace"), default="jump")
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--no-compress", action="store_true")
    ap.add_argument("--stages", type=int, choices=(1, 2), default=2)
    ap.add_argument("--jumps", type=int, default=1)
    ap.add_argument("--lanes", type=int, default=0, help="uneven workload: streams the groups run on "
                    "(0: one per group)")
    ap.add_argument("--tile", type=int, default=0, help="streamed tile shape (large workload): 0 64x16 "
                    "(64x32 lean tiles in the two-step forward), 1 128x8, 2 32x32, 3 64x32")
    ap.add_argument("--time-block", type=int, choices=(1, 2), default=2,
                    help="large workload: adjoint steps per streamed launch")
    ap.add_argument("--fwd-time-block", type=int, choices=(1, 2), default=2,
                    help="large workload: forward steps per streamed launch")
    ap.add_argument("--time-split", type=int, default=4,
                    help="default workload: time parts per unit of the resident work queue (0: chunk pipeline)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only; this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2
    if args.workload == "large":
        return bench_large(args)
    if args.workload == "copy":
        return bench_copy(args)
    if args.workload == "refine":
        return bench_refine(args)
    if args.workload == "uneven":
        return bench_uneven(args)
    if args.workload == "wavecache":
        return bench_wavecache(args)
    if args.workload == "physstage":
        return bench_physstage(args)
    if args.workload == "fwi":
        return bench_fwi(args)

    import ctypes

    import torch

    from wgnet_amd import _lib, synthetic
    from wgnet_amd.physics import WavePhysics

    B, S, nt, dt = args.models, 5, args.nt, 0.001
    ph = WavePhysics(shots=S, nt=nt, dt=dt)
    ph.engine.set_sources(SURFACE_SRC)
    _lib.check(ph.engine._lib.wg_set_time_split(ph.engine.handle, args.time_split))
    start, true = synthetic.batch(B, seed=1000)
    start_km = torch.from_numpy(start / 1000.0).cuda()
    obs = ph.simulate_torch(torch.from_numpy(true / 1000.0).cuda())
    fields = start_km.clone().requires_grad_(True)

    def step():
        losses = ph.misfit(fields, obs)
        losses.sum().backward()
        g = fields.grad
        fields.grad = None
        return losses, g

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
        losses, grads_km = step()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    lib.wg_timing_read(h, ctypes.byref(fms), ctypes.byref(fn), ctypes.byref(ams), ctypes.byref(an))
    lib.wg_timing_enable(h, 0)
    assert bool(torch.isfinite(losses).all())

    units = B * S
    value = units * args.steps / el
    # A timed region this short is over before a coarse utilisation sampler
    # looks: run the same step again, untimed, until 3 s of it have passed and
    # report that rate beside the timed one.
    sustained = None
    if el < 3.0:
        n_s, t1 = 0, time.perf_counter()
        while time.perf_counter() - t1 < 3.0:
            step()
            n_s += 1
            if n_s % 16 == 0:
                torch.cuda.synchronize()
        torch.cuda.synchronize()
        el_s = time.perf_counter() - t1
        sustained = {"steps": n_s, "seconds": round(el_s, 3), "value": round(units * n_s / el_s, 2)}
    # the last timed step's results, with the predicted gathers of the same
    # call (the queue is deterministic: the losses must repeat bit for bit),
    # for the oracle check of the CPU leg
    traces = torch.empty((B, S, nt, ph.engine.nrec), device="cuda")
    l_again, _ = ph.loss_and_grad_torch(fields.detach(), obs, traces_out=traces)
    repeat_equal = bool(torch.equal(l_again, losses.detach()))
    nx, nz, _ = ph.engine.dims()
    inner, R = (nx - 4) * (nz - 4), ph.engine.nrec
    # Algorithmic HBM bytes per unit: forward writes the snapshot (4 B per inner
    # cell-step) and the gathers (4 B per receiver-step); the adjoint reads the
    # snapshot and the residual.  A step may take several launches (chunks of
    # the overlapped pipeline): bytes per launch = bytes of the units of all
    # timed steps / launches.
    unit_bytes = 4 * inner * nt + 4 * R * nt
    fwd_bytes = unit_bytes * units * args.steps / max(fn.value, 1)
    adj_bytes = unit_bytes * units * args.steps / max(an.value, 1)
    f_avg = fms.value / max(fn.value, 1) / 1e3
    a_avg = ams.value / max(an.value, 1) / 1e3
    # the resident work queues (one forward and one adjoint launch per step)
    # or the chunk launches of the whole-unit kernels
    fk, ak = ("k_queue_fwd_fx", "k_queue_adj_fx") if args.time_split > 0 else ("k_forward_fx", "k_adjoint_fx")
    launch_ms = {fk: round(1e3 * f_avg, 3), ak: round(1e3 * a_avg, 3)}
    launches = {fk: fn.value / args.steps, ak: an.value / args.steps}
    if a_avg >= f_avg:
        kern, kbytes, kavg = ak, adj_bytes, a_avg
    else:
        kern, kbytes, kavg = fk, fwd_bytes, f_avg
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
            "time_split": args.time_split,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kern,
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "algorithmic_bytes_per_launch": int(kbytes),
            "launch_ms": launch_ms,
            "launches_per_step": launches,
            "shot_solve_hbm_frac": round(ss_bytes * value / 1e9 / HBM_PEAK_GBS, 4),
        },
    }
    if sustained:
        out["sustained"] = sustained
    run_cfg = {"models": B, "shots": S, "nt": nt, "time_split": args.time_split}
    traffic = profile_figure("traffic", kern, "bytes_per_launch", run_cfg)
    if traffic:
        out["roofline"]["traffic"] = traffic[0]
        out["roofline"]["traffic_source"] = traffic[1]
    valu = profile_figure("valu", kern, "sq_insts_valu_per_launch", run_cfg)
    if valu:
        # the kernel's real bound: VALU issue (a wave64 instruction takes 4 cycles
        # on a 16-lane SIMD; 256 CUs x 4 SIMDs at 2.4 GHz), not HBM: the profiled
        # instruction count of this workload over this run's launch time
        out["roofline"]["valu_issue_frac"] = round(valu[0] * 4 / (1024 * VALU_CLOCK_HZ * kavg), 4)
        out["roofline"]["valu_source"] = valu[1]
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        units_cpu = 64 * threads  # about 10-15 s of CPU work
        lv = losses.detach().cpu().numpy()
        models = list(dict.fromkeys([int(f + 3 * np.argmax(lv[f::3])) for f in range(3)] + [B - 1]))
        check = {"models": models, "obs": obs.cpu().numpy(), "losses": lv,
                 "grads_km": grads_km.cpu().numpy(), "traces": traces.cpu().numpy(),
                 "wavelet": ph.engine.wavelet(), "start": start, "true": true, "repeat_equal": repeat_equal}
        rate, cel, chk = cpu_baseline(start, true, nt, dt, threads, units_cpu, check)
        out["cpu_baseline"] = {
            "value": round(rate, 3),
            "unit": "shot-solves/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{units_cpu} (model, shot) forward+adjoint solves of the same workload, "
                      f"{threads} threads, {cel:.1f} s wall; float32 C restatement of the reference solver",
            **host_cpu(),
        }
        out["speedup_vs_cpu"] = round(value / rate, 1)
        out["oracle_check"] = chk
    if not args.no_refine:
        out["refine"] = refine_summary(args)
    print(json.dumps(out), flush=True)
    return 0


def refine_summary(args) -> dict:
    """The second half of the metric, refine solves-to-misfit, measured in the
    same run: the network trained in-run on solver pairs, then plain projected
    L-BFGS at the full budget (its final misfit is the target) and the
    learned arm at an eighth, a quarter and half of that budget."""
    r = run_refine(args, plain_fracs=(1.0,), learned_fracs=(0.125, 0.25, 0.5), pseudo=False,
                   other_modes=("first", "replace"), other_fracs=(0.5, 1.0))
    v = r["value"]
    return {
        "metric": "refine solves-to-misfit (learned / plain L-BFGS)",
        "value": v if np.isfinite(v) else None,
        "unit": "fraction of plain L-BFGS's solve budget to reach its final misfit (median over models; "
                "null = above one half)",
        "target": 0.5,
        "reached_at": r["reached_at"],
        "median_final_misfit": r["median_final_misfit"],
        "config": r["config"],
        "other_modes": r["other_modes"],
        "train": r["train"],
        "wall_s": r["wall_s"],
    }


def refine_scenes(n: int, seed: int):
    """Held-out refinement problems: dense = the starting model (km/s), true =
    start + anomaly, and a lattice of 7 x 7 active Gaussian slots (every fifth
    slot, sigma 1.5 px, amplitude +-0.05 km/s in a checkerboard)."""
    from wgnet_amd import synthetic
    from wgnet_amd.render import N_SLOTS, SLOT_SIDE

    start, true = synthetic.batch(n, seed=seed)
    g = np.zeros((n, N_SLOTS, 6), np.float32)
    for iz in range(2, SLOT_SIDE, 5):
        for ix in range(2, SLOT_SIDE, 5):
            k = iz * SLOT_SIDE + ix
            g[:, k, 2] = g[:, k, 3] = np.log(1.5)
            g[:, k, 5] = 0.05 if ((iz + ix) // 5) % 2 == 0 else -0.05
    return start / 1000.0, true / 1000.0, g


def refine_lattice_field():
    """[70, 70] km/s field that refine_scenes' initial active Gaussians render
    (the same for every scene): the scene's initial field is the starting
    model plus this."""
    import torch

    from wgnet_amd.render import ACTIVE_AMP_KM, render_residual

    lat = refine_scenes(1, seed=0)[2][0]
    act = np.flatnonzero(np.abs(lat[:, 5]) >= ACTIVE_AMP_KM)
    return render_residual(torch.from_numpy(lat[act])[None], torch.from_numpy(act))[0]


def run_refine(args, plain_fracs=(0.25, 0.5, 0.75, 1.0), learned_fracs=(0.125, 0.25, 0.5, 0.75, 1.0),
               pseudo: bool = True, other_modes=(), other_fracs=(0.5, 1.0)) -> dict:
    """Learned-preconditioned refinement against plain projected L-BFGS:
    solves needed to reach plain L-BFGS's final data misfit.  The network is
    trained in the run on pairs from the batched solver; plain L-BFGS runs at
    `plain_fracs` of the budget (its final misfit at the full budget is the
    target), the learned arm at `learned_fracs` and, with `pseudo`, the
    illumination-scaled arm at the plain fractions.  Each mode of
    `other_modes` ("first", "replace": the same network as a preconditioner
    of the L-BFGS direction instead of the --pc-mode arm) runs at
    `other_fracs`."""
    import torch

    from wgnet_amd.physics import WavePhysics
    from wgnet_amd.precond import Preconditioner, make_pairs, pseudo_hessian_hook, train
    from wgnet_amd.refine import refine_batch

    t_all = time.perf_counter()
    S, nt, dt = 5, 1000, 0.001
    bands = (6.0, 9.0, 12.0, 15.0)
    ph = WavePhysics(shots=S, nt=nt, dt=dt)
    ph.engine.set_sources(SURFACE_SRC)
    # stage 1: pairs at the scenes' initial fields (starting model + the
    # rendered initial Gaussians); stage 2 (--stages 2): pairs at the points
    # the stage-1 network's step reaches from the same models, trained
    # together with stage 1's; --train-pairs counts both
    n1 = args.train_pairs // args.stages
    offset = refine_lattice_field()
    net = Preconditioner(max_batch=256, seed=1)
    net.compress = not args.no_compress
    t_pairs = t_train = 0.0
    pairs = []
    for stage in range(args.stages):
        t0 = time.perf_counter()
        pairs.append(make_pairs(ph, n1, seed=4, bands_hz=bands, net=net if stage else None, offset_km=offset))
        torch.cuda.synchronize()
        t_pairs += time.perf_counter() - t0
        grads, fields, target = (torch.cat([p[k] for p in pairs]) for k in range(3))
        if stage == 0:
            baseline = float((target ** 2).mean())
        print(f"refine: stage {stage + 1}: {grads.shape[0]} training pairs (zero-prediction loss "
              f"{float((target ** 2).mean()):.4f})", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        hist = train(net, grads, fields, target, epochs=args.epochs, lr=args.lr, fit_scale=stage == 0)
        torch.cuda.synchronize()
        t_train += time.perf_counter() - t0
        print(f"refine: stage {stage + 1}: trained {args.epochs} epochs, loss {hist[0]:.4f} -> {hist[-1]:.4f}",
              file=sys.stderr, flush=True)
    del pairs
    if args.save_params:
        net.save(args.save_params)
    del grads, fields, target

    start_km, true_km, gauss = refine_scenes(args.eval_models, seed=900000)
    obs = ph.simulate_torch(torch.from_numpy(true_km).cuda())
    B = args.budget
    plain_final = {}
    t_plain = 0.0
    for f in plain_fracs:
        t0 = time.perf_counter()
        plain = refine_batch(start_km, gauss, obs, ph, budget_solves=int(B * f), bands_hz=bands)
        t_plain = time.perf_counter() - t0
        plain_final[f] = np.array([r.final_loss for r in plain])
        print(f"refine: plain L-BFGS {int(B * f)} solves in {t_plain:.1f} s, median final misfit "
              f"{np.median(plain_final[f]):.4e}", file=sys.stderr, flush=True)
    m_plain = plain_final[max(plain_fracs)]
    pc_final = {}
    if args.pc_mode == "jump":
        learned = {"update": net.update_hook(), "jumps_per_band": args.jumps}
    else:
        learned = {"precondition": net.hook(), "pc_mode": args.pc_mode}
    for f in learned_fracs:
        res = refine_batch(start_km, gauss, obs, ph, budget_solves=int(B * f), bands_hz=bands, **learned)
        pc_final[f] = np.array([r.final_loss for r in res])
        print(f"refine: learned preconditioner {int(B * f)} solves: median final misfit "
              f"{np.median(pc_final[f]):.4e}, reaches plain in {np.mean(pc_final[f] <= m_plain):.3f} of models",
              file=sys.stderr, flush=True)
    other_final = {}
    for mode in other_modes:
        if mode == args.pc_mode:
            continue
        kw = {"update": net.update_hook(), "jumps_per_band": args.jumps} if mode == "jump" else \
            {"precondition": net.hook(), "pc_mode": mode}
        other_final[mode] = {}
        for f in other_fracs:
            res = refine_batch(start_km, gauss, obs, ph, budget_solves=int(B * f), bands_hz=bands, **kw)
            other_final[mode][f] = np.array([r.final_loss for r in res])
            print(f"refine: network in {mode!r} mode, {int(B * f)} solves: median final misfit "
                  f"{np.median(other_final[mode][f]):.4e}, reaches plain in "
                  f"{np.mean(other_final[mode][f] <= m_plain):.3f} of models", file=sys.stderr, flush=True)
    phs_final = {}
    if pseudo:
        # the physics baseline: diagonal pseudo-Hessian (illumination) scaling;
        # its one extra forward per model is counted as one solve
        ph_hook = pseudo_hessian_hook(ph, torch.from_numpy(start_km).cuda())
        for f in plain_fracs:
            res = refine_batch(start_km, gauss, obs, ph, budget_solves=int(B * f) - 1, bands_hz=bands,
                               precondition=ph_hook, pc_mode="replace")
            phs_final[f] = np.array([r.final_loss for r in res])
            print(f"refine: pseudo-Hessian {int(B * f)} solves: median final misfit "
                  f"{np.median(phs_final[f]):.4e}, reaches plain in {np.mean(phs_final[f] <= m_plain):.3f} of "
                  f"models", file=sys.stderr, flush=True)

    def solves_needed(final):
        # the smallest tried fraction at which a model reaches plain's final
        # misfit (inf: none of them)
        need = np.full(len(m_plain), np.inf)
        for f in sorted(final, reverse=True):
            need[final[f] <= m_plain] = f
        return need

    need = solves_needed(pc_final)
    init = np.array([r.initial_loss for r in plain])
    out = {
        "metric": "refine solves-to-misfit: preconditioned solves to reach plain L-BFGS's final misfit / plain budget",
        "value": float(np.median(need)),
        "unit": "fraction of the plain budget (median over models; inf = not reached at the fractions tried)",
        "n_gpus": 1,
        "higher_is_better": False,
        "vs_baseline": None,
        "dtype": "f32 solver, bf16 network",
        "data": "synthetic (flat/curved/faulted 3-layer models + Gaussian anomaly)",
        "config": {"workload": f"{args.eval_models} held-out models, 5 shots, 70x70, nt={nt}, bands {bands} Hz, "
                               f"plain budget {B} solves; network trained on {args.train_pairs} solver pairs",
                   "budget": B, "eval_models": args.eval_models, "train_pairs": args.train_pairs,
                   "epochs": args.epochs, "stages": args.stages, "jumps_per_band": args.jumps,
                   "pc_mode": args.pc_mode, "compress": net.compress,
                   "layers": net.layers()},
        "fractions_tried": {"plain": list(plain_fracs), "learned": list(learned_fracs)},
        "reached_at": {str(f): float(np.mean(need <= f)) for f in learned_fracs},
        "median_final_misfit": {"initial": float(np.median(init)),
                                **{f"plain_{f}": float(np.median(plain_final[f])) for f in plain_fracs},
                                **{f"learned_{f}": float(np.median(pc_final[f])) for f in learned_fracs},
                                **{f"pseudo_hessian_{f}": float(np.median(phs_final[f])) for f in phs_final}},
        "learned_better_at_equal_budget": {str(f): float(np.mean(pc_final[f] < plain_final[f]))
                                           for f in learned_fracs if f in plain_final},
        "train": {"pairs_s": round(t_pairs, 2), "train_s": round(t_train, 2), "loss_first": hist[0],
                  "loss_last": hist[-1], "zero_prediction_loss": baseline, "grad_scale": net.grad_scale},
        "plain_refine_s": round(t_plain, 2),
        "wall_s": round(time.perf_counter() - t_all, 1),
    }
    out["other_modes"] = {}
    for mode, fin in other_final.items():
        need_m = solves_needed(fin)
        med = float(np.median(need_m))
        out["other_modes"][mode] = {
            "solves_needed_median": med if np.isfinite(med) else None,
            "reached_at": {str(f): float(np.mean(need_m <= f)) for f in fin},
            "median_final_misfit": {str(f): float(np.median(fin[f])) for f in fin}}
    if pseudo:
        need_ph = solves_needed(phs_final)
        out["pseudo_hessian"] = {"solves_needed_median": float(np.median(need_ph)),
                                 "reached_at": {str(f): float(np.mean(need_ph <= f)) for f in plain_fracs},
                                 "better_at_equal_budget": {str(f): float(np.mean(phs_final[f] < plain_final[f]))
                                                            for f in plain_fracs}}
    return out


def bench_refine(args) -> int:
    print(json.dumps(run_refine(args)), flush=True)
    return 0


def bench_uneven(args) -> int:
    """Uneven batch of solves (grid 70/100/141, nt 1000/1500/2000, 1-5 shots at
    random surface positions, passband 6/9/12/15 Hz) as one HIP graph."""
    import torch

    from wgnet_amd import synthetic
    from wgnet_amd.batch import UnevenBatch, random_solves

    solves = random_solves(args.solves, seed=2000)
    t0 = time.perf_counter()
    ub = UnevenBatch(solves, dt=0.001, lanes=args.lanes or None)
    tiles = ub.autotune(repeats=1)
    start, true = [], []
    for i, s in enumerate(solves):
        a, b = synthetic.batch(1, seed=2000 + 7919 * i, nvz=s.nvz, nvx=s.nvx)
        start.append(a[0])
        true.append(b[0])
    ub.set_models(true)
    ub.set_observed(ub.simulate(full_band=True))
    ub.set_models(start)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    print(f"uneven: {len(solves)} solves in {len(ub.groups)} (grid, nt) groups, tiles {tiles}, "
          f"set-up {t_setup:.1f} s", file=sys.stderr, flush=True)
    ub.capture()
    for _ in range(args.warmup):
        ub.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ub.replay()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    t1 = time.perf_counter()
    ub.run_eager()
    torch.cuda.synchronize()
    el_eager = time.perf_counter() - t1
    losses = torch.stack([l for l, _ in ub.results()])
    assert bool(torch.isfinite(losses).all())
    shot_solves = ub.active_shot_solves
    cell_steps = sum((s.nvx + 40) * (s.nvz + 40) * s.nt * s.shots for s in solves)
    value = shot_solves * args.steps / el
    def path(g):
        bands = g.engine._lib.wg_engine_bands(g.engine.handle)
        return f"resident, {bands} row bands per unit" if bands > 0 else (
            "streamed" if g.engine.streamed else "resident")

    groups = [{"grid": f"{g.key[0]}x{g.key[1]}", "nt": g.key[2], "solves": len(g.ids), "path": path(g)}
              for g in ub.groups]
    out = {
        "metric": "uneven batch fwd+adjoint shot-solves/s (grids 70/100/141, nt 1000-2000, 1-5 shots, bands)",
        "value": round(value, 2),
        "unit": "shot-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * el / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (flat/curved/faulted families per grid, Gaussian-anomaly observed data)",
        "config": {"workload": f"{len(solves)} solves, one HIP graph replay per step", "solves": len(solves),
                   "lanes": len(ub.lanes),
                   "shot_solves": shot_solves, "groups": groups, "tiles": {f"{k[0]}x{k[1]}x{k[2]}": v
                                                                            for k, v in tiles.items()}},
        "inner_gcell_steps_per_s": round(2 * cell_steps * args.steps / el / 1e9, 2),
        "eager_ms": round(1e3 * el_eager, 2),
    }
    print(json.dumps(out), flush=True)
    return 0


def bench_wavecache(args) -> int:
    """The training-data cache writer: `cache-models` synthetic 70 x 70 models
    x 5 shots, 2001 steps of 0.5 ms, decimated to 251 samples, float16 file
    with progress and RMS bookkeeping, in batches of `cache-batch` models.
    A step is the whole cache (fresh directory each step); the reference's
    call pattern (one engine.simulate per model) is timed beside it on the
    same engine."""
    import ctypes
    import shutil
    import tempfile

    import torch

    from wgnet_amd import synthetic, wavecache

    n, S, nt, dt = args.cache_models, 5, 2001, 0.0005
    phys = wavecache.make_physics()
    _, true = synthetic.batch(n, seed=4000)
    fields = {"train": (true / 1000.0).astype(np.float32)}
    root = tempfile.mkdtemp(prefix="wavecache_", dir=os.environ.get("TMPDIR", "/tmp"))
    lib, h = phys.engine._lib, phys.engine.handle
    fms, fn, ams, an = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double(), ctypes.c_int64()

    def step(k):
        d = os.path.join(root, f"s{k}")
        wavecache.ensure_wave_cache(fields, d, ("train",), phys, batch=args.cache_batch, log=lambda *_: None)
        return d

    try:
        for k in range(args.warmup):
            shutil.rmtree(step(-1 - k))
        torch.cuda.synchronize()
        lib.wg_timing_enable(h, 1)
        lib.wg_timing_read(h, ctypes.byref(fms), ctypes.byref(fn), ctypes.byref(ams), ctypes.byref(an))
        t0 = time.perf_counter()
        for k in range(args.steps):
            d = step(k)
            print(f"wavecache: step {k} done", file=sys.stderr, flush=True)
            if k + 1 < args.steps:
                shutil.rmtree(d)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        lib.wg_timing_read(h, ctypes.byref(fms), ctypes.byref(fn), ctypes.byref(ams), ctypes.byref(an))
        lib.wg_timing_enable(h, 0)
        with open(os.path.join(d, "waves_train.done.json")) as f:
            rms = json.load(f)["rms"]
        assert np.isfinite(rms) and rms > 0.0
        # the reference's call pattern: one host-buffer simulate per model
        m_loop = min(args.cache_loop_models, n)
        t1 = time.perf_counter()
        for i in range(m_loop):
            phys.engine.simulate(np.ascontiguousarray(fields["train"][i].reshape(-1) * 1000.0,
                                                      dtype=np.float32))[:, ::8, :]
        el_loop = time.perf_counter() - t1
    finally:
        shutil.rmtree(root, ignore_errors=True)

    value = n * args.steps / el
    nx, nz, _ = phys.engine.dims()
    inner, R = (nx - 4) * (nz - 4), phys.engine.nrec
    f_avg = fms.value / max(fn.value, 1) / 1e3
    units_per_launch = n * S * args.steps / max(fn.value, 1)
    # Algorithmic HBM bytes of the non-storing forward: the gathers written
    # (4 B per receiver-step of each unit) and each model's velocity read once
    # (4 B per cell; its shots share it).
    f_bytes = int(units_per_launch * 4 * R * nt + units_per_launch / S * 4 * phys.engine.nvx * phys.engine.nvz)
    achieved = f_bytes / f_avg / 1e9
    kern = "k_forward_fx_nostore"
    out = {
        "metric": "wave-cache models/s (70x70, 5 shots, 2001 steps of 0.5 ms, decimated float16 cache + RMS)",
        "value": round(value, 2),
        "unit": "models/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * el / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (flat/curved/faulted 3-layer models)",
        "config": {"workload": f"{n} models x {S} shots, 70x70 grid (114x114 padded), nt={nt}, dt=0.5 ms, "
                               f"batches of {args.cache_batch} models, waves_train.f16.bin + progress/done json",
                   "models": n, "shots": S, "nt": nt, "batch": args.cache_batch},
        "roofline": {
            "bound": "hbm",
            "kernel": kern,
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": None,
            "algorithmic_bytes_per_launch": f_bytes,
            "launch_ms": {kern: round(1e3 * f_avg, 3)},
            "launches_per_step": fn.value / args.steps,
        },
        "per_model_loop": {"models": m_loop, "models_per_s": round(m_loop / el_loop, 2) if m_loop else None,
                           "note": "engine.simulate of one model at a time (data.py:160-163), same engine"},
        "gpu_forward_share": round(fms.value / 1e3 / el, 3),
    }
    run_cfg = {"workload": "wavecache", "batch": args.cache_batch, "shots": S, "nt": nt}
    traffic = profile_figure("traffic", kern, "bytes_per_launch", run_cfg)
    if traffic:
        out["roofline"]["traffic"], out["roofline"]["traffic_source"] = traffic
    valu = profile_figure("valu", kern, "sq_insts_valu_per_launch", run_cfg)
    if valu:
        out["roofline"]["valu_issue_frac"] = round(valu[0] * 4 / (1024 * VALU_CLOCK_HZ * f_avg), 4)
        out["roofline"]["valu_source"] = valu[1]
    if not args.no_cpu:
        from oracle import oracle as O

        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        m_cpu = 24 * threads  # whole models, about 10-15 s of CPU work
        g = O.Geometry(nt=nt, dt=dt)
        w = O.ricker_f32(nt, dt, 15.0)
        src = np.tile(phys.engine.sources(), (m_cpu, 1))
        vel = np.ascontiguousarray(np.repeat(true[:m_cpu].reshape(m_cpu, -1), S, 0))
        O.batch_forward_f32(g, src[:threads], vel[:threads], w, threads)  # warm-up
        t2 = time.perf_counter()
        traces = O.batch_forward_f32(g, src, vel, w, threads).reshape(m_cpu, S, nt, R)[:, :, ::8, :]
        sq = 0.0
        for i in range(m_cpu):
            sq += float(np.square(traces[i], dtype=np.float64).sum())
            traces[i].astype(np.float16)
        cel = time.perf_counter() - t2
        rate = m_cpu / cel
        out["cpu_baseline"] = {
            "value": round(rate, 3),
            "unit": "models/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{m_cpu} models x {S} shots simulated at nt={nt} ({m_cpu * S} forward solves), decimated, "
                      f"float16 cast and square-summed, {threads} threads, {cel:.1f} s wall; "
                      "float32 C restatement of the reference solver",
        }
        out["speedup_vs_cpu"] = round(value / rate, 1)
    print(json.dumps(out), flush=True)
    return 0


def bench_physstage(args) -> int:
    """The wave-equation term of the phys fine-tuning stage at the reference's
    settings (5 shots, 2001 steps of 0.5 ms, phys_batch models, phys_ratio 0.5,
    bands 6/9/12/full over the timed steps).  A step: re-simulate the observed
    data of phys_batch true models, misfit and adjoint gradient of the
    predicted fields, gradient-norm balancing and the backward through a
    stand-in for the network's output (field = start + 0.2 tanh(w), L1
    supervised loss).  The reference's per-model host loop on the same engine
    is timed beside it."""
    import ctypes

    import torch

    from wgnet_amd import _lib, synthetic
    from wgnet_amd.physics import WavePhysics
    from wgnet_amd.physstage import PhysStage, band_at

    nb, S, nt, B = args.phys_batch, 5, 2001, max(32, args.phys_batch)
    dev = torch.device("cuda")
    total = args.warmup + args.steps
    stage = PhysStage(steps=total, bands="6,9,12,full", phys_ratio=0.5, phys_batch=nb)
    eng = stage.physics.engine
    _lib.check(eng._lib.wg_set_time_split(eng.handle, args.time_split))
    start, true = synthetic.batch(B, seed=5000)
    base = torch.from_numpy(start / 1000.0).to(dev)
    true_km = torch.from_numpy((true / 1000.0).astype(np.float32)).to(dev)
    w = torch.zeros(B, 70, 70, device=dev, requires_grad=True)
    opt = torch.optim.SGD([w], lr=1.0)

    def step(k):
        field = base + 0.2 * torch.tanh(w)
        loss = (field - true_km).abs().mean()
        loss, parts = stage(loss, field, true_km, k)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        return loss, parts

    for k in range(1, args.warmup + 1):
        step(k)
    torch.cuda.synchronize()
    lib, h = eng._lib, eng.handle
    fms, fn, ams, an = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double(), ctypes.c_int64()
    lib.wg_timing_enable(h, 1)
    lib.wg_timing_read(h, ctypes.byref(fms), ctypes.byref(fn), ctypes.byref(ams), ctypes.byref(an))
    t0 = time.perf_counter()
    bands = []
    for k in range(args.warmup + 1, total + 1):
        loss, parts = step(k)
        bands.append(parts["band_hz"])
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    lib.wg_timing_read(h, ctypes.byref(fms), ctypes.byref(fn), ctypes.byref(ams), ctypes.byref(an))
    lib.wg_timing_enable(h, 0)
    assert bool(torch.isfinite(loss))

    # the reference's pattern: per-model host simulate and evaluate_adjoint
    ref = WavePhysics()
    t1 = time.perf_counter()
    for k in range(args.ref_steps):
        ref.set_band(band_at(k + 1, args.ref_steps, stage.bands))
        field = (base + 0.2 * torch.tanh(w)).detach()[:nb].cpu().numpy()
        tn = true_km[:nb].cpu().numpy()
        obs = [ref.engine.simulate(np.ascontiguousarray(tn[b].reshape(-1) * 1000.0, dtype=np.float32))
               for b in range(nb)]
        for b in range(nb):
            ref.engine.evaluate_adjoint(np.ascontiguousarray(field[b].reshape(-1) * 1000.0, dtype=np.float32),
                                        np.ascontiguousarray(obs[b].reshape(-1)))
    el_ref = (time.perf_counter() - t1) / max(1, args.ref_steps)

    value = nb * args.steps / el
    nx, nz, _ = eng.dims()
    inner, R = (nx - 4) * (nz - 4), eng.nrec
    units = nb * S
    unit_bytes = 4 * inner * nt + 4 * R * nt
    # forward launches: the observed simulation (no snapshots) and the storing
    # forward; the adjoint launches read the snapshots and the residual
    adj_bytes = unit_bytes * units * args.steps / max(an.value, 1)
    a_avg = ams.value / max(an.value, 1) / 1e3
    f_avg = fms.value / max(fn.value, 1) / 1e3
    # the work queues run only when there are more units than compute units
    cus = torch.cuda.get_device_properties(dev).multi_processor_count
    ak = "k_queue_adj_fx" if args.time_split > 0 and units > cus else "k_adjoint_fx"
    achieved = adj_bytes / a_avg / 1e9
    out = {
        "metric": "phys-stage wave term models/s (5 shots, 2001 steps of 0.5 ms, observed re-simulated, balanced "
                  "adjoint gradient)",
        "value": round(value, 2),
        "unit": "models/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * el / args.steps, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (flat/curved/faulted 3-layer models, Gaussian-anomaly truth); stand-in network output",
        "config": {"workload": f"phys_batch {nb} models x {S} shots per training step, 70x70 (114x114 padded), "
                               f"nt={nt}, dt=0.5 ms, phys_ratio 0.5, bands 6/9/12/full",
                   "phys_batch": nb, "shots": S, "nt": nt, "bands_timed": sorted(set(bands)),
                   "time_split": args.time_split},
        "roofline": {
            "bound": "hbm",
            "kernel": ak,
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": None,
            "algorithmic_bytes_per_launch": int(adj_bytes),
            "launch_ms": {"k_forward_fx (observed and storing, side by side), mean": round(1e3 * f_avg, 3),
                          ak: round(1e3 * a_avg, 3)},
            "launches_per_step": {"forward": fn.value / args.steps, "adjoint": an.value / args.steps},
        },
        "per_model_loop": {"ms_per_step": round(1e3 * el_ref, 2), "steps": args.ref_steps,
                           "note": "physics.py:59-76's per-model engine.simulate and evaluate_adjoint through host "
                                   "arrays, same engine"},
        "speedup_vs_per_model_loop": round(el_ref / (el / args.steps), 2),
    }
    if not args.no_cpu:
        from oracle import oracle as O

        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        g = O.Geometry(nt=nt, dt=0.0005)
        wav = eng.wavelet()
        src = np.tile(eng.sources(), (nb, 1))
        tv = np.ascontiguousarray(np.repeat(true[:nb].reshape(nb, -1), S, 0))
        cv = np.ascontiguousarray(np.repeat(start[:nb].reshape(nb, -1), S, 0))
        reps = 0
        t2 = time.perf_counter()
        while reps < 20 and (reps == 0 or time.perf_counter() - t2 < 10.0):  # about 10 s of CPU work
            obs = O.batch_forward_f32(g, src, tv, wav, threads)
            O.batch_fwd_adj_f32(g, src, cv, wav, obs, threads)
            reps += 1
        cel = time.perf_counter() - t2
        out["cpu_baseline"] = {
            "value": round(nb * reps / cel, 3),
            "unit": "models/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{reps} steps of {nb} models x {S} shots, observed forward plus forward+adjoint at nt={nt} "
                      f"({2 * nb * S} solves a step), {threads} threads, {cel:.1f} s wall; float32 C restatement of "
                      "the reference solver",
        }
        out["speedup_vs_cpu"] = round(value / (nb * reps / cel), 1)
    print(json.dumps(out), flush=True)
    return 0


def bench_fwi(args) -> int:
    """Budgeted multiscale pixel-grid FWI (L-BFGS-B, bounds [1.4, 4.6] km/s,
    bands 5/8/12/full, gradient smoothing 1.5 cells) from two starts per
    model, the training-set mean 1D profile and the macro model (truth
    low-passed at 6 cells), at the engine defaults (5 shots, 2001 steps of
    0.5 ms): fwi-models x 2 problems in lockstep, one batched engine call per
    round.  One problem through `fwi` (one model per solve, as the reference
    runs them) is timed beside it."""
    import ctypes

    import torch

    from wgnet_amd import synthetic
    from wgnet_amd.fwi_eval import _default_workers, fwi, fwi_batch, macro_start
    from wgnet_amd.physics import WavePhysics

    M, S, nt, bands = args.fwi_models, 5, 2001, [5.0, 8.0, 12.0, None]
    phys = WavePhysics()
    start, true = synthetic.batch(M, seed=6000)
    truth = (true / 1000.0).astype(np.float32)
    linear1d = np.repeat((start / 1000.0).mean(axis=(0, 2))[None, :, None], 70, axis=2).astype(np.float32)
    starts = np.concatenate([np.repeat(linear1d, M, 0), macro_start(truth, 6.0)])
    tm = np.concatenate([truth, truth])
    observed = phys.simulate(tm)
    P = starts.shape[0]
    lib, h = phys.engine._lib, phys.engine.handle
    fms, fn, ams, an = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double(), ctypes.c_int64()
    torch.cuda.synchronize()
    lib.wg_timing_enable(h, 1)
    lib.wg_timing_read(h, ctypes.byref(fms), ctypes.byref(fn), ctypes.byref(ams), ctypes.byref(an))
    t0 = time.perf_counter()
    finals, hists, calls = fwi_batch(starts, observed, phys, bands, args.fwi_budget)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    lib.wg_timing_read(h, ctypes.byref(fms), ctypes.byref(fn), ctypes.byref(ams), ctypes.byref(an))
    lib.wg_timing_enable(h, 0)
    print(f"fwi: {P} problems in {el:.1f} s, {calls} batched calls", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    fwi(starts[0], observed[0], phys, bands, args.fwi_budget)
    el_one = time.perf_counter() - t1

    solves = [sum(x["solves"] for x in hh) for hh in hists]
    mae0 = np.abs(starts - tm).mean(axis=(1, 2)) * 1000.0
    mae1 = np.abs(np.stack(finals) - tm).mean(axis=(1, 2)) * 1000.0
    value = P / el
    nx, nz, _ = phys.engine.dims()
    inner, R = (nx - 4) * (nz - 4), phys.engine.nrec
    unit_bytes = 4 * inner * nt + 4 * R * nt
    a_avg = ams.value / max(an.value, 1) / 1e3
    f_avg = fms.value / max(fn.value, 1) / 1e3
    adj_bytes = unit_bytes * S * sum(solves) / max(an.value, 1)
    out = {
        "metric": "pixel-grid FWI problems/s (L-BFGS-B, 5/8/12/full bands, 5 shots, 2001 steps of 0.5 ms)",
        "value": round(value, 3),
        "unit": "problems/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(1e3 * el, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (flat/curved/faulted 3-layer models, Gaussian-anomaly truth)",
        "config": {"workload": f"{P} FWI problems ({M} models x linear-1D and macro starts), budget "
                               f"{args.fwi_budget} solves over 4 bands, lockstep batches",
                   "problems": P, "budget": args.fwi_budget, "bands": "5,8,12,full", "smooth_grad": 1.5,
                   "host_workers": _default_workers(P)},
        "solves": {"total": int(sum(solves)), "per_problem_mean": round(float(np.mean(solves)), 1),
                   "batched_calls": calls, "solves_per_s": round(sum(solves) / el, 1)},
        "mae_mps": {"linear1d": [round(float(mae0[:M].mean()), 1), round(float(mae1[:M].mean()), 1)],
                    "macro": [round(float(mae0[M:].mean()), 1), round(float(mae1[M:].mean()), 1)]},
        "roofline": {
            "bound": "hbm",
            "kernel": "k_adjoint_fx",
            "achieved": round(adj_bytes / a_avg / 1e9, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(adj_bytes / a_avg / 1e9 / HBM_PEAK_GBS, 5),
            "traffic": None,
            "algorithmic_bytes_per_launch": int(adj_bytes),
            "launch_ms": {"k_forward_fx": round(1e3 * f_avg, 3), "k_adjoint_fx": round(1e3 * a_avg, 3)},
        },
        "one_problem_fwi_s": round(el_one, 2),
        "speedup_vs_one_at_a_time": round(el_one * P / el, 1),
    }
    if not args.no_cpu:
        from oracle import oracle as O

        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        g = O.Geometry(nt=nt, dt=0.0005)
        phys.set_band(None)
        wav = phys.engine.wavelet()
        m_cpu = max(1, 2 * threads // S)
        src = np.tile(phys.engine.sources(), (m_cpu, 1))
        cv = np.ascontiguousarray(np.repeat(starts[:m_cpu].reshape(m_cpu, -1) * 1000.0, S, 0), dtype=np.float32)
        obs = np.ascontiguousarray(observed[:m_cpu].reshape(m_cpu * S, nt, R))
        reps = 0
        t2 = time.perf_counter()
        while reps < 20 and (reps == 0 or time.perf_counter() - t2 < 10.0):  # about 10 s of CPU work
            O.batch_fwd_adj_f32(g, src, cv, wav, obs, threads)
            reps += 1
        cel = time.perf_counter() - t2
        rate = m_cpu * reps / cel / float(np.mean(solves))
        out["cpu_baseline"] = {
            "value": round(rate, 5),
            "unit": "problems/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{reps} rounds of {m_cpu} models x {S} shots forward+adjoint at nt={nt}, {threads} threads, "
                      f"{cel:.1f} s wall, divided by the {float(np.mean(solves)):.1f} solves a problem takes; float32 "
                      "C restatement of the reference solver (the optimiser's host work not counted)",
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
    eng.set_time_block(args.time_block)
    eng.set_forward_time_block(args.fwd_time_block)
    eng.set_stream_tile(args.tile)
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
    adj_spl = nt * args.steps / max(an.value, 1)  # adjoint steps per launch (2 with temporal blocking)
    adj_kern = "k_adj_step2" if adj_spl > 1.5 else "k_adj_step"
    fwd_spl = 2 * nt * args.steps / max(fn.value, 1)  # forward and recompute: steps per launch
    fwd_kern = "k_fwd_step2" if fwd_spl > 1.5 else "k_fwd_step"
    adj_bytes = int(20 * inner * units * adj_spl)
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
            "kernel": adj_kern,
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "algorithmic_bytes_per_launch": adj_bytes,
            "launch_ms": {fwd_kern: round(1e3 * f_avg, 4), adj_kern: round(1e3 * a_avg, 4)},
            "adjoint_steps_per_launch": round(adj_spl, 3),
            "forward_steps_per_launch": round(fwd_spl, 3),
            "shot_solve_hbm_frac": round(ss_bytes * value / 1e9 / HBM_PEAK_GBS, 4),
        },
        "dot_test": {"operator": "wavelet -> traces, 1 model x 20 shots, full nt", "rel": dot_rel},
    }
    run_cfg = {"workload": "large", "models": B, "shots": S, "grid": f"{nvx}x{nvz}", "tile": args.tile,
               "time_block": args.time_block}
    traffic = profile_figure("traffic", adj_kern, "bytes_per_launch", run_cfg)
    if traffic:
        out["roofline"]["traffic"], out["roofline"]["traffic_source"] = traffic
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        nt_cpu = 500
        units_cpu = 16 * threads  # about 25 s of CPU work
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
