"""Budgeted test-time refinement of a velocity scene against observed gathers.

Counterpart of python/wgnet/refine.py (:20-169) over the HIP engine.  The
scene is  field = dense + bilinear(coarse 9x9) + render(active Gaussians)
(refine.py:83-90).  Projected L-BFGS (two-loop recursion, Nocedal & Wright
Alg. 7.4, with H0 = s'y / y'y and 8 pairs) with Armijo backtracking (c = 1e-4,
at most 6 halvings, step norm <= 1), a frequency-continuation schedule over
passbands, and the reference's solve accounting: 1 solve per misfit-only
evaluation, 2 per evaluation with gradient (refine.py:99-108).

Each model's optimiser is a generator (`_program`) that yields evaluation
requests and receives their results; `refine` drives one program, and
`refine_batch` drives many in lockstep, answering every round's requests with
one batched misfit call and one batched misfit-and-gradient call in which each
model carries its own passband.  A model's sequence of evaluations and
decisions is the same either way.

`precondition`, when given, maps the velocity-space gradient (the hook point of
refine.py:105-106) and the current field, both km/s-based [n, 70, 70] device
tensors (and, if it takes a third argument, the models' indices in the batch),
to a velocity-space vector that the chain rule carries to the search
parameters.  Two ways of using it (`pc_mode`):
  "first"    L-BFGS runs on plain gradients; the first step of each band goes
             along the preconditioned direction (rescaled to the gradient's
             norm) when that is a descent direction of the true misfit, and
             is judged by the Armijo test with the true slope.
  "replace"  L-BFGS runs on the preconditioned gradients; a model whose
             preconditioned line search fails to reduce the misfit continues
             with plain gradients (its L-BFGS memory is reset).

`update`, when given, maps the same inputs to a predicted model update (km/s,
[n, 70, 70]).  It is fitted onto the scene parameters (least squares between
the scene's field change and the update; no solves) and the first
`jumps_per_band` steps of each band try the fitted parameter step at full
length, each from the point the previous one reached, judged by the Armijo
test with the true misfit slope.  A learned step (`update`, or `precondition` in
"first" mode) that fails its line search costs its trial solves and the band
continues with L-BFGS.
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
    preconditioned: bool = False  # still preconditioned at the end


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


class _Group:
    """Models that share one set of active Gaussian slots: their scenes are
    assembled, and chain-ruled, together."""

    def __init__(self, dense_km: np.ndarray, active: np.ndarray, device):
        self.device = torch.device(device)
        self.active = active
        self.n_g = len(active)
        self.slots = torch.as_tensor(active, device=self.device)
        self.dense = torch.as_tensor(np.asarray(dense_km, np.float32), device=self.device)  # [m, 70, 70]

    def assemble(self, rows: list[int], vecs: list[np.ndarray], need_grad: bool):
        """fields [n, 70, 70] of members `rows` at parameter vectors `vecs`,
        and the leaf tensor of those vectors."""
        v = torch.tensor(np.stack(vecs), dtype=torch.float32, device=self.device, requires_grad=need_grad)
        return self.field_of(rows, v), v

    def field_of(self, rows: list[int], v: torch.Tensor) -> torch.Tensor:
        """fields [n, 70, 70] of members `rows` at parameters v [n, 81 + 6 K]."""
        n, n_c = len(rows), COARSE * COARSE
        coarse = v[:, :n_c].reshape(n, 1, COARSE, COARSE)
        field = self.dense[rows] + F.interpolate(coarse, size=(GRID, GRID), mode="bilinear", align_corners=True)[:, 0]
        if self.n_g:
            field = field + render_residual(v[:, n_c:].reshape(n, self.n_g, 6), self.slots)
        return field


class _Scene:
    """One model's scene: parameters are [coarse 81 | 6 per active Gaussian]
    in scaled units (refine.py:62-90); `group`/`row` locate its dense field."""

    def __init__(self, group: _Group, row: int, gaussians: np.ndarray):
        self.group, self.row = group, row
        self.device = group.device
        self.active = group.active
        self.n_g = group.n_g
        self.n_c = COARSE * COARSE
        self.slots = group.slots
        self.dense_t = group.dense[row]
        self.scales = np.concatenate([np.full(self.n_c, COARSE_SCALE), np.tile(np.array(GAUSS_SCALES), self.n_g)])
        self.x0 = np.concatenate([np.zeros(self.n_c), gaussians[self.active].astype(np.float64).reshape(-1)])

    def project(self, vec: np.ndarray) -> np.ndarray:
        out = vec.copy()
        n_c, n_g = self.n_c, self.n_g
        out[:n_c] = np.clip(out[:n_c], -COARSE_MAX_KM, COARSE_MAX_KM)
        if n_g:
            out[n_c:] = clamp_gaussians(torch.from_numpy(out[n_c:].reshape(n_g, 6))).numpy().reshape(-1)
        return out

    def field(self, vec: np.ndarray) -> np.ndarray:
        with torch.no_grad():
            return self.group.assemble([self.row], [vec], False)[0][0].cpu().numpy()


class _Request(NamedTuple):
    need_grad: bool
    vec: np.ndarray
    band: float | None
    precondition: bool
    jump: bool = False


def _program(scene: _Scene, budget_solves: int, bands_hz: Sequence[float | None], use_pc: bool,
             pc_mode: str = "first", jumps_per_band: int = 0):
    """One model's projected L-BFGS with the band schedule (refine.py:92-169).
    Yields _Request; receives (loss, plain_grad, pc_grad, jump_step) for
    gradient requests (pc_grad / jump_step None when not asked for; all in
    the scene's scaled units) and the loss otherwise.
    Returns (x, initial_loss, final_loss, solves, history, preconditioned)."""
    solves = 0
    history: list = []
    x = scene.project(scene.x0)
    initial_loss = yield _Request(False, x, None, False)
    solves += 1
    per_band = budget_solves // len(bands_hz)
    scales = scene.scales
    first = use_pc and pc_mode == "first"
    if first:
        use_pc = False

    for band in bands_hz:
        band_budget = solves + per_band
        jumps_left = jumps_per_band
        cur_loss, g_plain, g_pc, g_jump = yield _Request(True, x, band, use_pc or first, jumps_left > 0)
        solves += 2
        cur_grad = g_pc if use_pc else g_plain
        s_hist: list = []
        y_hist: list = []
        rho: list = []
        learned = first and g_pc is not None
        while solves + 1 < band_budget:
            d = lbfgs_direction(cur_grad, s_hist, y_hist, rho)
            slope = float(cur_grad @ d)
            took_learned = took_jump = False
            if jumps_left > 0 and g_jump is not None:
                # the band's first steps try the fitted predicted update, at
                # full length, judged by the true misfit slope
                jumps_left -= 1
                sl = float(g_plain @ g_jump)
                if np.all(np.isfinite(g_jump)) and np.isfinite(sl) and sl < 0.0:
                    d, slope = g_jump, sl
                    took_jump = True
            elif learned:
                # the band's first step goes along the learned direction,
                # judged by the true misfit slope
                learned = False
                dl = -g_pc * (np.linalg.norm(cur_grad) / max(float(np.linalg.norm(g_pc)), 1e-300))
                sl = float(cur_grad @ dl)
                if np.all(np.isfinite(dl)) and np.isfinite(sl) and sl < 0.0:
                    d, slope = dl, sl
                    took_learned = True
            if not np.isfinite(slope) or slope >= 0.0:
                s_hist, y_hist, rho = [], [], []
                d = -cur_grad
                slope = -float(cur_grad @ cur_grad)
            norm = float(np.linalg.norm(d))
            if norm > 1.0 and not took_jump:
                d, slope = d / norm, slope / norm
            step, accepted = 1.0, False
            for _ in range(LINE_SEARCH_TRIALS):
                if solves + 1 >= band_budget:
                    break
                cand = scene.project(x + step * scales * d)
                cand_loss = yield _Request(False, cand, band, False)
                solves += 1
                if cand_loss <= cur_loss + ARMIJO * step * slope:
                    if solves + 2 > band_budget:
                        x, cur_loss = cand, cand_loss
                        accepted = True
                        break
                    cand_loss, c_plain, c_pc, g_jump = yield _Request(True, cand, band, use_pc, jumps_left > 0)
                    solves += 2
                    cand_grad = c_pc if use_pc else c_plain
                    s, y = (cand - x) / scales, cand_grad - cur_grad
                    sy = float(s @ y)
                    if sy > 1e-12:
                        if len(s_hist) >= LBFGS_PAIRS:
                            s_hist.pop(0), y_hist.pop(0), rho.pop(0)
                        s_hist.append(s), y_hist.append(y), rho.append(1.0 / sy)
                    x, cur_loss, cur_grad, g_plain = cand, cand_loss, cand_grad, c_plain
                    accepted = True
                    break
                step *= 0.5
            history.append({"band_hz": band, "solves": solves, "loss": cur_loss, "accepted": accepted,
                            "preconditioned": use_pc or took_learned or took_jump, "jump": took_jump})
            if not accepted:
                if took_jump or took_learned:
                    jumps_left = 0
                    continue  # L-BFGS from the same point
                if use_pc:
                    # plain L-BFGS from here on: the last plain gradient is at x
                    use_pc = False
                    cur_grad = g_plain
                    s_hist, y_hist, rho = [], [], []
                    continue
                break

    final_loss = yield _Request(False, x, None, False)
    solves += 1
    return x, initial_loss, final_loss, solves, history, use_pc


def _takes_ids(fn) -> bool:
    """Whether a precondition hook takes a third argument: the indices of
    the models (in the refined batch) whose gradients it receives."""
    import inspect

    try:
        params = inspect.signature(fn).parameters.values()
    except (TypeError, ValueError):
        return False
    return sum(p.kind in (p.POSITIONAL_ONLY, p.POSITIONAL_OR_KEYWORD) for p in params) >= 3 or any(
        p.kind == p.VAR_POSITIONAL for p in params)


def _misfit_f64(pred: torch.Tensor, observed: torch.Tensor) -> torch.Tensor:
    """Per-model 0.5 ||p - d||^2 / ||d||^2 accumulated in float64."""
    d = (pred.double() - observed.double()).flatten(1)
    o = observed.double().flatten(1)
    return 0.5 * (d * d).sum(1) / (o * o).sum(1)


def _fields(scenes, pending, ids, need_grad):
    """Fields [n, 70, 70] of the requests `ids` (in that order), assembled per
    group; returns them with the (group, leaf, positions) needed to chain-rule."""
    by_group: dict = {}
    for pos, i in enumerate(ids):
        by_group.setdefault(id(scenes[i].group), []).append(pos)
    out = [None] * len(ids)
    parts = []
    for positions in by_group.values():
        grp = scenes[ids[positions[0]]].group
        f, leaf = grp.assemble([scenes[ids[p]].row for p in positions], [pending[ids[p]].vec for p in positions],
                               need_grad)
        for k, p in enumerate(positions):
            out[p] = f[k]
        parts.append((f, leaf, positions))
    return torch.stack(out), parts


def _chain(field_t, parts, gv: torch.Tensor, retain: bool) -> list[np.ndarray]:
    """Per request, the parameter gradient (unscaled) of <field, gv>."""
    res = [None] * gv.shape[0]
    for f, leaf, positions in parts:
        w = gv[torch.tensor(positions, device=gv.device)].to(f.device)
        (g,) = torch.autograd.grad((f * w).sum(), [leaf], retain_graph=retain)
        g = g.detach().cpu().numpy().astype(np.float64)
        for k, p in enumerate(positions):
            res[p] = g[k]
    return res


def _fit_update(scenes, pending, ids, update: torch.Tensor, iters: int = 40) -> list[np.ndarray]:
    """Per request `ids[k]`, the parameter step u (unscaled) whose field change
    best matches the velocity update `update[k]` (km/s) in least squares:
    min_u ||field(x + u) - field(x) - update||^2, solved per group by L-BFGS
    in the scene's scaled units (u = scales * delta), no solves."""
    out = [None] * len(ids)
    by_group: dict = {}
    for pos, i in enumerate(ids):
        by_group.setdefault(id(scenes[i].group), []).append(pos)
    for positions in by_group.values():
        grp = scenes[ids[positions[0]]].group
        rows = [scenes[ids[p]].row for p in positions]
        x = torch.tensor(np.stack([pending[ids[p]].vec for p in positions]), dtype=torch.float32, device=grp.device)
        sc = torch.tensor(np.stack([scenes[ids[p]].scales for p in positions]), dtype=torch.float32,
                          device=grp.device)
        with torch.no_grad():
            want = grp.field_of(rows, x) + update[torch.tensor(positions, device=update.device)].to(grp.device)
        delta = torch.zeros_like(x, requires_grad=True)
        opt = torch.optim.LBFGS([delta], lr=1.0, max_iter=iters, history_size=10, line_search_fn="strong_wolfe")

        def closure():
            opt.zero_grad()
            loss = ((grp.field_of(rows, x + sc * delta) - want) ** 2).sum()
            loss.backward()
            return loss

        opt.step(closure)
        u = (sc * delta).detach().cpu().numpy().astype(np.float64)
        for k, p in enumerate(positions):
            out[p] = u[k]
    return out


def _drive(scenes, programs, physics: WavePhysics, observed_full: torch.Tensor, precondition, update=None):
    """Runs the programs in lockstep; returns their return values."""
    dev = physics.device
    band_obs: dict = {}
    band_wav: dict = {}

    def observed(i, band):
        key = (i, band)
        if key not in band_obs:
            band_obs[key] = physics.bandlimit_to(observed_full[i:i + 1], band)[0].contiguous()
        return band_obs[key]

    def wavelet(band):
        if band not in band_wav:
            band_wav[band] = torch.from_numpy(physics.band_wavelet(band)).to(dev)
        return band_wav[band]

    pending = {i: next(p) for i, p in enumerate(programs)}
    results: dict = {}
    while pending:
        answers = {}
        loss_ids = [i for i, r in pending.items() if not r.need_grad]
        grad_ids = [i for i, r in pending.items() if r.need_grad]
        if loss_ids:
            with torch.no_grad():
                fields = _fields(scenes, pending, loss_ids, False)[0].to(dev)
            wav = torch.stack([wavelet(pending[i].band) for i in loss_ids])
            pred = physics.simulate_torch(fields, wavelets=wav)
            obs = torch.stack([observed(i, pending[i].band) for i in loss_ids])
            losses = _misfit_f64(pred, obs).cpu().numpy()
            for k, i in enumerate(loss_ids):
                answers[i] = float(losses[k])
        if grad_ids:
            field_t, parts = _fields(scenes, pending, grad_ids, True)
            fields = field_t.detach().to(dev)
            wav = torch.stack([wavelet(pending[i].band) for i in grad_ids])
            obs = torch.stack([observed(i, pending[i].band) for i in grad_ids])
            losses, grads = physics.loss_and_grad_torch(fields, obs, wavelets=wav)
            pc_rows = [k for k, i in enumerate(grad_ids) if pending[i].precondition]
            pcg = None
            if pc_rows and precondition is not None:
                sel = torch.tensor(pc_rows, device=dev)
                if _takes_ids(precondition):
                    pcg = precondition(grads[sel], fields[sel], [grad_ids[k] for k in pc_rows])
                else:
                    pcg = precondition(grads[sel], fields[sel])
            jumps = [None] * len(grad_ids)
            jump_rows = [k for k, i in enumerate(grad_ids) if pending[i].jump]
            if jump_rows and update is not None:
                sel = torch.tensor(jump_rows, device=dev)
                ids = [grad_ids[k] for k in jump_rows]
                upd = update(grads[sel], fields[sel], ids) if _takes_ids(update) else update(grads[sel], fields[sel])
                for k, u in zip(jump_rows, _fit_update(scenes, pending, ids, upd.detach().float())):
                    jumps[k] = u / scenes[grad_ids[k]].scales
            losses = losses.cpu().numpy()
            plain = _chain(field_t, parts, grads, retain=pcg is not None)
            pc_full = None
            if pcg is not None:
                gv = torch.zeros_like(grads)
                gv[torch.tensor(pc_rows, device=dev)] = pcg
                pc_full = _chain(field_t, parts, gv, retain=False)
            for k, i in enumerate(grad_ids):
                want_pc = pending[i].precondition and pcg is not None
                sc = scenes[i].scales
                answers[i] = (float(losses[k]), plain[k] * sc, pc_full[k] * sc if want_pc else None, jumps[k])
        for i, ans in answers.items():
            try:
                pending[i] = programs[i].send(ans)
            except StopIteration as stop:
                results[i] = stop.value
                del pending[i]
    return [results[i] for i in range(len(programs))]


def _result(scene: _Scene, out, seconds: float) -> RefineResult:
    x, initial_loss, final_loss, solves, history, pc = out
    initial_field = scene.field(scene.project(scene.x0))
    final_field = scene.field(x)
    n_c, n_g = scene.n_c, scene.n_g
    return RefineResult(
        initial_field_km=initial_field,
        final_field_km=final_field,
        initial_loss=initial_loss,
        final_loss=final_loss,
        solves=solves,
        seconds=seconds,
        history=history,
        gaussians=x[n_c:].reshape(n_g, 6).astype(np.float32),
        coarse_km=x[:n_c].reshape(COARSE, COARSE).astype(np.float32),
        preconditioned=pc,
    )


def refine_batch(
    dense_km: np.ndarray,
    gaussians: np.ndarray,
    observed_full,
    physics: WavePhysics,
    budget_solves: int = 160,
    bands_hz: Sequence[float | None] = (8.0, 12.0, None),
    amp_keep_km: float = ACTIVE_AMP_KM,
    device: torch.device | str = "cpu",
    precondition: Callable[[torch.Tensor, torch.Tensor], torch.Tensor] | None = None,
    pc_mode: str = "first",
    update: Callable[[torch.Tensor, torch.Tensor], torch.Tensor] | None = None,
    jumps_per_band: int = 1,
) -> list[RefineResult]:
    """Refines B scenes together: dense_km [B, 70, 70], gaussians [B, 1225, 6]
    activated, observed_full [B, shots, nt, 70] (numpy or tensor).  Scene
    algebra runs on `device`; the solves on the physics device."""
    t0 = time.perf_counter()
    obs = observed_full if isinstance(observed_full, torch.Tensor) else torch.from_numpy(
        np.ascontiguousarray(observed_full, np.float32))
    obs = obs.to(physics.device, torch.float32)
    B = len(dense_km)
    actives = [np.flatnonzero(np.abs(gaussians[b][:, 5]) >= amp_keep_km) for b in range(B)]
    groups: dict = {}
    for b in range(B):
        groups.setdefault(actives[b].tobytes(), []).append(b)
    scenes: list = [None] * B
    for members in groups.values():
        grp = _Group(np.stack([dense_km[b] for b in members]), actives[members[0]], device)
        for row, b in enumerate(members):
            scenes[b] = _Scene(grp, row, gaussians[b])
    if pc_mode not in ("first", "replace"):
        raise ValueError("pc_mode must be 'first' or 'replace'")
    if update is not None and precondition is not None and pc_mode == "first":
        raise ValueError("update and a 'first'-mode precondition both claim each band's first step")
    jumps = int(jumps_per_band) if update is not None else 0
    programs = [_program(sc, budget_solves, tuple(bands_hz), precondition is not None, pc_mode, jumps)
                for sc in scenes]
    outs = _drive(scenes, programs, physics, obs, precondition, update)
    el = time.perf_counter() - t0
    return [_result(sc, out, el) for sc, out in zip(scenes, outs)]


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
    observed_full [shots, nt, 70] (refine.py:56-169).  `precondition` here maps
    numpy [70, 70] (gradient, field) to numpy [70, 70]."""
    hook = None
    if precondition is not None:
        def hook(g: torch.Tensor, f: torch.Tensor) -> torch.Tensor:
            out = [precondition(g[k].cpu().numpy(), f[k].cpu().numpy()) for k in range(g.shape[0])]
            return torch.from_numpy(np.stack(out).astype(np.float32)).to(g.device)
    return refine_batch(np.asarray(dense_km)[None], np.asarray(gaussians)[None], np.asarray(observed_full)[None],
                        physics, budget_solves, bands_hz, amp_keep_km, device, hook)[0]
