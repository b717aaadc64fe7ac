"""Budgeted multiscale pixel-grid FWI over the batched engine.

Counterpart of `fwi` / `misfit_full` (python/wgnet/fwi_eval.py:29-63) and of
the macro-model start it compares against (igm.py:52-63).  The optimiser is
the reference's: scipy's L-BFGS-B on the 70 x 70 pixel grid in km/s, bounds
[1.4, 4.6], maxfun = maxiter = max(2, budget // bands) per band, maxcor 10,
the exact adjoint gradient Gaussian-smoothed by `smooth_grad` cells
(scipy.ndimage, mode "nearest"), the observed data band-limited per band by
the physics' causal low-pass.

`fwi` runs one problem as the reference does.  `fwi_batch` runs many in
lockstep: each problem's L-BFGS-B runs unchanged in its own thread, and
whenever every running problem waits for an objective value, their models are
evaluated in one batched engine call, each with its own passband wavelet
(problems move through the bands at their own pace).  The optimiser side is
host work that the interpreter lock serialises, so by default the problems
are shared among worker processes (`_fwi_host`, numpy and scipy only), and
this process, which owns the GPU, evaluates every round's requests from all
of them in one call.  Per-model results of the batched call equal one-model
calls bit for bit, so each problem follows exactly the trajectory that `fwi`
gives it alone.
"""

from __future__ import annotations

import math
import os
import subprocess
import sys

import numpy as np
import torch

from ._fwi_host import N, V_BOUNDS_KM, lbfgsb_bands, recv, run_lockstep, send

MAX_WORKERS = 16


def fwi(start_km: np.ndarray, observed_full: np.ndarray, physics, bands: list, budget: int,
        smooth_grad: float = 1.5) -> tuple[np.ndarray, list[dict]]:
    """fwi_eval.py:32-57: one problem, one model per solve.  Returns the final
    model [70, 70] km/s and the per-band history (band_hz, misfit, solves)."""

    def band_observed(band):
        physics.set_band(band)
        return physics.bandlimit(np.asarray(observed_full, np.float32))

    def evaluate(field, band, observed):
        losses, grads = physics.loss_and_grad(field[None], observed[None])
        return losses[0], grads[0]

    out = lbfgsb_bands(start_km, bands, budget, smooth_grad, band_observed, evaluate)
    physics.set_band(None)
    return out


def misfit_full(model_km: np.ndarray, observed_full: np.ndarray, physics) -> float:
    """fwi_eval.py:60-63: the full-band misfit of one model."""
    physics.set_band(None)
    losses, _ = physics.loss_and_grad(np.asarray(model_km, np.float32)[None], np.asarray(observed_full)[None])
    return float(losses[0])


class _Evaluator:
    """One batched loss-and-gradient call for the requests of a round, each
    model in its own passband against its own band-limited observed data
    (physics.bandlimit's CPU filter, as `fwi` applies it, kept on the device
    while the problem stays in that band)."""

    def __init__(self, physics, full: np.ndarray):
        self.physics, self.full = physics, full
        self.wavelets: dict = {}
        self.observed: dict = {}  # pid -> (band, device gathers)
        self.calls = 0

    def wavelet(self, band):
        if band not in self.wavelets:
            w = self.physics.band_wavelet(band)
            self.wavelets[band] = torch.from_numpy(np.ascontiguousarray(w, np.float32)).to(self.physics.device)
        return self.wavelets[band]

    def band_observed(self, pid: int, band):
        cur = self.observed.get(pid)
        if cur is None or cur[0] != band:
            d = np.ascontiguousarray(self.physics.bandlimit_to(self.full[pid], band), np.float32)
            cur = (band, torch.from_numpy(d).to(self.physics.device))
            self.observed[pid] = cur
        return cur[1]

    def __call__(self, requests: list) -> dict:
        """requests: [(pid, band, field [70, 70] float32)] -> {pid: (loss, grad)}."""
        dev = self.physics.device
        fields = torch.from_numpy(np.stack([f for _, _, f in requests])).to(dev)
        obs = torch.stack([self.band_observed(pid, band) for pid, band, _ in requests])
        wav = torch.stack([self.wavelet(band) for _, band, _ in requests])
        losses, grads = self.physics.loss_and_grad_torch(fields, obs, wavelets=wav)
        losses, grads = losses.cpu().numpy(), grads.cpu().numpy()
        self.calls += 1
        return {pid: (losses[k], grads[k]) for k, (pid, _, _) in enumerate(requests)}


def _default_workers(P: int) -> int:
    try:
        cpus = len(os.sched_getaffinity(0))
    except AttributeError:
        cpus = os.cpu_count() or 1
    return max(1, min(P, MAX_WORKERS, cpus))


def _run_workers(starts, bands, budget, smooth_grad, ev: _Evaluator, W: int):
    """Problems shared among W worker processes; one batched call per round
    for the requests of all of them."""
    P = starts.shape[0]
    shares = [s for s in np.array_split(np.arange(P), W) if len(s)]
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env = dict(os.environ)
    env["PYTHONPATH"] = root + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
    for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS"):
        env[k] = "1"
    procs = [subprocess.Popen([sys.executable, "-m", "wgnet_amd._fwi_host"], stdin=subprocess.PIPE,
                              stdout=subprocess.PIPE, env=env, cwd=root) for _ in shares]
    results: dict = {}
    try:
        for pr, share in zip(procs, shares):
            send(pr.stdin, {"pids": [int(p) for p in share], "starts": [starts[p] for p in share],
                            "bands": list(bands), "budget": int(budget), "smooth_grad": float(smooth_grad)})
        live = list(range(len(procs)))
        while live:
            waiting, requests = [], []
            for w in live:
                kind, body = recv(procs[w].stdout)
                if kind == "eval":
                    waiting.append(w)
                    requests.append(body)
                elif kind == "done":
                    results.update(body)
                else:
                    raise RuntimeError(f"FWI worker failed:\n{body}")
            live = waiting
            if requests:
                answers = ev([r for req in requests for r in req])
                for w, req in zip(waiting, requests):
                    send(procs[w].stdin, {pid: answers[pid] for pid, _, _ in req})
    finally:
        for pr in procs:
            try:
                pr.stdin.close()
            except OSError:
                pass
        for pr in procs:
            try:
                pr.wait(timeout=10)
            except subprocess.TimeoutExpired:
                pr.kill()
                pr.wait()
    return results


def fwi_batch(starts_km: np.ndarray, observed_full: np.ndarray, physics, bands: list, budget: int,
              smooth_grad: float = 1.5, workers: int | None = None) -> tuple[list[np.ndarray], list[list[dict]], int]:
    """`fwi` for P problems at once: starts_km [P, 70, 70] km/s, observed_full
    [P, shots, nt, 70] full-band gathers.  `workers` worker processes run the
    optimiser side (default: one per host core, at most 16 and at most P;
    0 runs every problem as a thread of this process).  Returns the P final
    models, their histories (each equal to what `fwi` returns for that
    problem alone) and the number of batched engine calls made."""
    starts = np.asarray(starts_km, np.float32)
    full = np.asarray(observed_full, np.float32)
    P = starts.shape[0]
    if full.shape[0] != P:
        raise ValueError("one observed gather set per start")
    ev = _Evaluator(physics, full)
    W = _default_workers(P) if workers is None else int(workers)
    if W <= 0:
        res, _ = run_lockstep(list(range(P)), list(starts), bands, budget, smooth_grad, ev.band_observed,
                              lambda pending: ev([(pid, band, f) for pid, (f, band, _) in sorted(pending.items())]))
    else:
        res = _run_workers(starts, bands, budget, smooth_grad, ev, W)
    return [res[p][0] for p in range(P)], [res[p][1] for p in range(P)], ev.calls


def macro_start(v_km: np.ndarray, sigma_cells: float) -> np.ndarray:
    """The macro-model start of fwi_eval.py:106 (igm.py:52-63): a separable
    Gaussian low-pass of [B, 70, 70] km/s, taps exp(-t^2 / 2 sigma^2) for
    |t| <= ceil(3 sigma) normalised to unit sum, reflecting at the edges
    without repeating the edge cell; along x, then along z, in float32."""
    v = np.asarray(v_km, np.float32)
    if sigma_cells <= 0:
        return v.copy()
    half = int(math.ceil(3.0 * sigma_cells))
    t = np.arange(-half, half + 1, dtype=np.float32)
    taps = np.exp(np.float32(-0.5) * (t / np.float32(sigma_cells)) ** 2).astype(np.float32)
    taps /= taps.sum(dtype=np.float32)

    def along(a: np.ndarray, axis: int) -> np.ndarray:
        pad = [(0, 0)] * a.ndim
        pad[axis] = (half, half)
        ap = np.pad(a, pad, mode="reflect")
        n = a.shape[axis]
        out = np.zeros_like(a)
        for j, w in enumerate(taps):
            out += w * np.take(ap, np.arange(j, j + n), axis=axis)
        return out

    return along(along(v, v.ndim - 1), v.ndim - 2)


__all__ = ["V_BOUNDS_KM", "fwi", "fwi_batch", "misfit_full", "macro_start"]
