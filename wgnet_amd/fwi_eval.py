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
(problems move through the bands at their own pace).  Per-model results of
the batched call equal one-model calls bit for bit, so each problem follows
exactly the trajectory that `fwi` gives it alone.
"""

from __future__ import annotations

import math
import threading

import numpy as np
import torch
from scipy.ndimage import gaussian_filter
from scipy.optimize import minimize

from .physics import N

V_BOUNDS_KM = (1.4, 4.6)  # fwi_eval.py:29


def _lbfgsb_bands(start_km, observed_full, bands, budget, smooth_grad, band_observed, evaluate):
    """The band loop of fwi_eval.py:37-57.  band_observed(band) -> a handle of
    the band-limited observed data; evaluate(field_f32 [N, N], band, handle)
    -> (loss, grad [N, N] float32)."""
    x = start_km.astype(np.float64).reshape(-1)
    per_band = max(2, budget // len(bands))
    history = []
    for band in bands:
        observed = band_observed(band)
        calls = 0

        def objective(vec: np.ndarray):
            nonlocal calls
            calls += 1
            loss, grad = evaluate(vec.reshape(N, N).astype(np.float32), band, observed)
            g = gaussian_filter(grad, smooth_grad, mode="nearest") if smooth_grad > 0 else grad
            return float(loss), g.reshape(-1).astype(np.float64)

        res = minimize(objective, x, jac=True, method="L-BFGS-B", bounds=[V_BOUNDS_KM] * (N * N),
                       options={"maxfun": per_band, "maxiter": per_band, "maxcor": 10})
        x = res.x
        history.append({"band_hz": band, "misfit": float(res.fun), "solves": calls})
    return x.reshape(N, N).astype(np.float32), history


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

    out = _lbfgsb_bands(start_km, observed_full, bands, budget, smooth_grad, band_observed, evaluate)
    physics.set_band(None)
    return out


def misfit_full(model_km: np.ndarray, observed_full: np.ndarray, physics) -> float:
    """fwi_eval.py:60-63: the full-band misfit of one model."""
    physics.set_band(None)
    losses, _ = physics.loss_and_grad(np.asarray(model_km, np.float32)[None], np.asarray(observed_full)[None])
    return float(losses[0])


class _Lockstep:
    """Collects one objective request per running problem and answers them
    with one batched loss-and-gradient call."""

    def __init__(self, physics, problems: int):
        self.physics = physics
        self.cond = threading.Condition()
        self.active = problems
        self.pending: dict[int, tuple] = {}
        self.results: dict[int, tuple] = {}
        self.wavelets: dict = {}
        self.calls = 0
        self.error: BaseException | None = None

    def wavelet(self, band):
        if band not in self.wavelets:
            w = self.physics.band_wavelet(band)
            self.wavelets[band] = torch.from_numpy(np.ascontiguousarray(w, np.float32)).to(self.physics.device)
        return self.wavelets[band]

    def _run_batch(self) -> None:
        ids = sorted(self.pending)
        try:
            fields = torch.from_numpy(np.stack([self.pending[i][0] for i in ids])).to(self.physics.device)
            obs = torch.stack([self.pending[i][2] for i in ids])
            wav = torch.stack([self.wavelet(self.pending[i][1]) for i in ids])
            losses, grads = self.physics.loss_and_grad_torch(fields, obs, wavelets=wav)
            losses, grads = losses.cpu().numpy(), grads.cpu().numpy()
            for k, i in enumerate(ids):
                self.results[i] = (losses[k], grads[k])
        except BaseException as exc:  # every waiting problem sees the failure
            self.error = exc
        self.calls += 1
        self.pending.clear()
        self.cond.notify_all()

    def submit(self, pid: int, field: np.ndarray, band, observed: torch.Tensor):
        with self.cond:
            if self.error is not None:
                raise RuntimeError("a batched evaluation failed") from self.error
            self.pending[pid] = (field, band, observed)
            if len(self.pending) == self.active:
                self._run_batch()
            while pid not in self.results and self.error is None:
                self.cond.wait()
            if pid not in self.results:
                raise RuntimeError("a batched evaluation failed") from self.error
            return self.results.pop(pid)

    def finish(self) -> None:
        with self.cond:
            self.active -= 1
            if self.pending and len(self.pending) == self.active:
                self._run_batch()


def fwi_batch(starts_km: np.ndarray, observed_full: np.ndarray, physics, bands: list, budget: int,
              smooth_grad: float = 1.5) -> tuple[list[np.ndarray], list[list[dict]], int]:
    """`fwi` for P problems at once: starts_km [P, 70, 70] km/s, observed_full
    [P, shots, nt, 70] full-band gathers.  Returns the P final models, their
    histories (each equal to what `fwi` returns for that problem alone) and
    the number of batched engine calls made."""
    starts = np.asarray(starts_km, np.float32)
    full = np.asarray(observed_full, np.float32)
    P = starts.shape[0]
    if full.shape[0] != P:
        raise ValueError("one observed gather set per start")
    hub = _Lockstep(physics, P)
    finals: list = [None] * P
    histories: list = [None] * P
    errors: list = [None] * P

    def run(p: int) -> None:
        def band_observed(band):
            # physics.bandlimit's CPU filter, as `fwi` applies it, then on the device
            return torch.from_numpy(np.ascontiguousarray(physics.bandlimit_to(full[p], band),
                                                         np.float32)).to(physics.device)

        try:
            finals[p], histories[p] = _lbfgsb_bands(starts[p], full[p], bands, budget, smooth_grad, band_observed,
                                                    lambda f, band, obs: hub.submit(p, f, band, obs))
        except BaseException as exc:
            errors[p] = exc
        finally:
            hub.finish()

    threads = [threading.Thread(target=run, args=(p,), daemon=True) for p in range(P)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for e in errors:
        if e is not None:
            raise e
    return finals, histories, hub.calls


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
