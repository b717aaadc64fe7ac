"""Budgeted test-time refinement of a velocity scene against observed gathers.

Counterpart of python/wgnet/refine.py (:20-169) over the HIP engine.  The
scene is  field = dense + bilinear(coarse 9x9) + render(active Gaussians)
(refine.py:83-90).  Projected L-BFGS (two-loop recursion, Nocedal & Wright
Alg. 7.4, with H0 = s'y / y'y and 8 pairs) with Armijo backtracking (c = 1e-4,
at most 6 halvings, step norm <= 1), a frequency-continuation schedule over
passbands, and the reference's solve accounting: 1 solve per misfit-only
evaluation, 2 per evaluation with gradient (refine.py:99-108).

`precondition`, when given, maps the velocity-space gradient [70, 70] (the
hook point of refine.py:105-106) and the current field [70, 70] to the
velocity-space vector that the chain rule carries to the search parameters.
"""

from __future__ import annotations

import time
from typing import Callable, NamedTuple, Sequence

import numpy as np
import torch
import torch.nn.functional as F

from .physics import WavePhysics
from .render import ACTIVE_AMP_KM, GRID, clamp_gaussians, render_residual

COARSE = 9
COARSE_MAX_KM = 1.0
GAUSS_SCALES = (1.0, 1.0, 0.2, 0.2, 0.2, 0.1)
COARSE_SCALE = 0.1
LBFGS_PAIRS = 8
LINE_SEARCH_TRIALS = 6
ARMIJO = 1e-4


class RefineResult(NamedTuple):
    initial_field_km: np.ndarray
    final_field_km: np.ndarray
    initial_loss: float
    final_loss: float
    solves: int
    seconds: float
    history: list
    gaussians: np.ndarray  # [K, 6] refined active slots
    coarse_km: np.ndarray  # [9, 9]


def lbfgs_direction(grad: np.ndarray, s_hist: list, y_hist: list, rho: list) -> np.ndarray:
    """-H grad by the two-loop recursion (refine.py:41-53)."""
    q = grad.copy()
    alphas = []
    for s, y, r in zip(reversed(s_hist), reversed(y_hist), reversed(rho)):
        a = r * float(s @ q)
        alphas.append(a)
        q -= a * y
    if s_hist:
        q *= float(s_hist[-1] @ y_hist[-1]) / max(float(y_hist[-1] @ y_hist[-1]), 1e-30)
    for s, y, r, a in zip(s_hist, y_hist, rho, reversed(alphas)):
        q += s * (a - r * float(y @ q))
    return -q


def _misfit_f64(pred: np.ndarray, observed: np.ndarray) -> float:
    d = pred.reshape(-1).astype(np.float64) - observed.reshape(-1).astype(np.float64)
    o = observed.reshape(-1).astype(np.float64)
    return 0.5 * float(d @ d) / float(o @ o)


def refine(
    dense_km: np.ndarray,
    gaussians: np.ndarray,
    observed_full: np.ndarray,
    physics: WavePhysics,
    budget_solves: int = 160,
    bands_hz: Sequence[float | None] = (8.0, 12.0, None),
    amp_keep_km: float = ACTIVE_AMP_KM,
    device: torch.device | str = "cpu",
    precondition: Callable[[np.ndarray, np.ndarray], np.ndarray] | None = None,
) -> RefineResult:
    """Refine one scene.  dense_km [70, 70], gaussians [1225, 6] activated,
    observed_full [shots, nt, 70] (refine.py:56-169)."""
    t0 = time.perf_counter()
    active = np.flatnonzero(np.abs(gaussians[:, 5]) >= amp_keep_km)
    n_g = len(active)
    slots = torch.as_tensor(active, device=device)
    dense_t = torch.as_tensor(np.asarray(dense_km, np.float32), device=device)
    n_c = COARSE * COARSE
    scales = np.concatenate([np.full(n_c, COARSE_SCALE), np.tile(np.array(GAUSS_SCALES), n_g)])
    x = np.concatenate([np.zeros(n_c), gaussians[active].astype(np.float64).reshape(-1)])

    def project(vec: np.ndarray) -> np.ndarray:
        out = vec.copy()
        out[:n_c] = np.clip(out[:n_c], -COARSE_MAX_KM, COARSE_MAX_KM)
        if n_g:
            out[n_c:] = clamp_gaussians(torch.from_numpy(out[n_c:].reshape(n_g, 6))).numpy().reshape(-1)
        return out

    def assemble(vec: np.ndarray, need_grad: bool):
        coarse = torch.tensor(vec[:n_c].reshape(1, 1, COARSE, COARSE), dtype=torch.float32, device=device,
                              requires_grad=need_grad)
        g = torch.tensor(vec[n_c:].reshape(n_g, 6), dtype=torch.float32, device=device, requires_grad=need_grad)
        field = dense_t + F.interpolate(coarse, size=(GRID, GRID), mode="bilinear", align_corners=True)[0, 0]
        if n_g:
            field = field + render_residual(g[None], slots)[0]
        return field, coarse, g

    solves = 0
    history: list = []

    def evaluate(vec: np.ndarray, need_grad: bool, observed: np.ndarray):
        nonlocal solves
        field, coarse, g = assemble(vec, need_grad)
        field_np = field.detach().cpu().numpy()
        if not need_grad:
            solves += 1
            pred = physics.engine.simulate(np.ascontiguousarray(field_np.reshape(-1) * 1000.0, dtype=np.float32))
            return _misfit_f64(pred, observed), field_np, None
        solves += 2
        losses, grads = physics.loss_and_grad(field_np[None], observed[None])
        gv = grads[0]
        if precondition is not None:
            gv = precondition(gv, field_np)
        gc, gg = torch.autograd.grad((field * torch.as_tensor(gv, device=device)).sum(), [coarse, g],
                                     allow_unused=True)
        parts = [gc.detach().cpu().numpy().reshape(-1)]
        parts.append(gg.detach().cpu().numpy().reshape(-1) if gg is not None else np.zeros(6 * n_g))
        return float(losses[0]), field_np, np.concatenate(parts).astype(np.float64) * scales

    x = project(x)
    physics.set_band(None)
    initial_loss, initial_field, _ = evaluate(x, False, observed_full)
    per_band = budget_solves // len(bands_hz)

    for band in bands_hz:
        physics.set_band(band)
        observed = physics.bandlimit(observed_full)
        band_budget = solves + per_band
        cur_loss, _, cur_grad = evaluate(x, True, observed)
        s_hist: list = []
        y_hist: list = []
        rho: list = []
        while solves + 1 < band_budget:
            d = lbfgs_direction(cur_grad, s_hist, y_hist, rho)
            slope = float(cur_grad @ d)
            if not np.isfinite(slope) or slope >= 0.0:
                s_hist, y_hist, rho = [], [], []
                d = -cur_grad
                slope = -float(cur_grad @ cur_grad)
            norm = float(np.linalg.norm(d))
            if norm > 1.0:
                d, slope = d / norm, slope / norm
            step, accepted = 1.0, False
            for _ in range(LINE_SEARCH_TRIALS):
                if solves + 1 >= band_budget:
                    break
                cand = project(x + step * scales * d)
                cand_loss, _, _ = evaluate(cand, False, observed)
                if cand_loss <= cur_loss + ARMIJO * step * slope:
                    if solves + 2 > band_budget:
                        x, cur_loss = cand, cand_loss
                        accepted = True
                        break
                    cand_loss, _, cand_grad = evaluate(cand, True, observed)
                    s, y = (cand - x) / scales, cand_grad - cur_grad
                    sy = float(s @ y)
                    if sy > 1e-12:
                        if len(s_hist) >= LBFGS_PAIRS:
                            s_hist.pop(0), y_hist.pop(0), rho.pop(0)
                        s_hist.append(s), y_hist.append(y), rho.append(1.0 / sy)
                    x, cur_loss, cur_grad = cand, cand_loss, cand_grad
                    accepted = True
                    break
                step *= 0.5
            history.append({"band_hz": band, "solves": solves, "loss": cur_loss, "accepted": accepted})
            if not accepted:
                break

    physics.set_band(None)
    final_loss, final_field, _ = evaluate(x, False, observed_full)
    return RefineResult(
        initial_field_km=initial_field,
        final_field_km=final_field,
        initial_loss=initial_loss,
        final_loss=final_loss,
        solves=solves,
        seconds=time.perf_counter() - t0,
        history=history,
        gaussians=x[n_c:].reshape(n_g, 6).astype(np.float32),
        coarse_km=x[:n_c].reshape(COARSE, COARSE).astype(np.float32),
    )
