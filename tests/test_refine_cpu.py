"""Renderer and L-BFGS pieces of refine() on the CPU (render.py:86-109,
refine.py:41-53; verify.py:42-54 for the renderer VJP)."""

from __future__ import annotations

import math

import numpy as np
import torch

from wgnet_amd import render as R
from wgnet_amd.refine import lbfgs_direction


def random_scene(gen: torch.Generator, n_active: int = 40) -> torch.Tensor:
    """verify.py:34-39."""
    raw = torch.randn(1, R.N_SLOTS, 6, generator=gen) * 0.5
    raw[..., 5] = 0.0
    idx = torch.randperm(R.N_SLOTS, generator=gen)[:n_active]
    raw[0, idx, 5] = torch.randn(n_active, generator=gen) * 1.5
    return R.activate_gaussians(raw)


def test_renderer_vjp_fd():
    """verify.py:42-54: autograd VJP vs central FD in float64, 1e-5."""
    gen = torch.Generator().manual_seed(1)
    params = random_scene(gen).double().requires_grad_(True)
    cot = torch.randn(1, 70, 70, generator=gen).double()
    (g,) = torch.autograd.grad((R.render_residual(params) * cot).sum(), params)
    direction = torch.randn_like(params) * torch.tensor([0.3, 0.3, 0.1, 0.1, 0.1, 0.1]).double()
    eps = 1e-4
    with torch.no_grad():
        fp = (R.render_residual(params + eps * direction) * cot).sum()
        fm = (R.render_residual(params - eps * direction) * cot).sum()
    fd, adj = float((fp - fm) / (2 * eps)), float((g * direction).sum())
    assert abs(fd - adj) / max(abs(fd), 1e-9) < 1e-5


def test_renderer_single_gaussian_closed_form():
    q = torch.zeros(1, 1, 6, dtype=torch.float64)
    q[0, 0, 0], q[0, 0, 1] = 0.4, -0.7
    q[0, 0, 2], q[0, 0, 3], q[0, 0, 4], q[0, 0, 5] = math.log(2.0), math.log(3.0), 0.6, 0.7
    iz, ix = 10, 12
    f = R.render_residual(q, torch.tensor([iz * R.SLOT_SIDE + ix]))[0]
    z, x = torch.meshgrid(torch.arange(70.0, dtype=torch.float64), torch.arange(70.0, dtype=torch.float64), indexing="ij")
    dx, dz = (x + 0.5) - (2 * ix + 1 + 0.4), (z + 0.5) - (2 * iz + 1 - 0.7)
    u = (dx * math.cos(0.6) + dz * math.sin(0.6)) / 2.0
    v = (dz * math.cos(0.6) - dx * math.sin(0.6)) / 3.0
    ref = 0.7 * torch.exp(-0.5 * (u * u + v * v))
    window = ((z - 2 * iz).abs() <= R.WINDOW) & ((x - 2 * ix).abs() <= R.WINDOW)
    assert float((f - ref * window).abs().max()) < 1e-6
    # a subset of slots renders what the full slot set renders with the others at zero amplitude
    full = torch.zeros(1, R.N_SLOTS, 6, dtype=torch.float64)
    full[0, iz * R.SLOT_SIDE + ix] = q[0, 0]
    assert float((R.render_residual(full)[0] - f).abs().max()) < 1e-12


def test_clamp_and_activate_bounds():
    p = R.activate_gaussians(torch.randn(1000, 6) * 10)
    assert torch.equal(R.clamp_gaussians(p), p)
    c = R.clamp_gaussians(torch.tensor([[5.0, -5.0, -3.0, 3.0, 4.0, -2.0]]))
    np.testing.assert_allclose(c[0].numpy(), [2, -2, math.log(0.5), math.log(6), math.pi / 2, -1], rtol=1e-6)


def _bfgs_matrix(s_hist, y_hist, n):
    """Textbook inverse-Hessian recursion H <- (I - r s y')H(I - r y s') + r s s'
    from H0 = (s'y / y'y) I of the newest pair."""
    s, y = s_hist[-1], y_hist[-1]
    H = np.eye(n) * float(s @ y) / float(y @ y)
    for s, y in zip(s_hist, y_hist):
        r = 1.0 / float(s @ y)
        V = np.eye(n) - r * np.outer(y, s)
        H = V.T @ H @ V + r * np.outer(s, s)
    return H


def test_lbfgs_two_loop_equals_bfgs_matrix():
    rng = np.random.default_rng(0)
    n = 7
    A = rng.standard_normal((n, n))
    A = A @ A.T + n * np.eye(n)
    s_hist = [rng.standard_normal(n) for _ in range(4)]
    y_hist = [A @ s for s in s_hist]
    rho = [1.0 / float(s @ y) for s, y in zip(s_hist, y_hist)]
    g = rng.standard_normal(n)
    d = lbfgs_direction(g, s_hist, y_hist, rho)
    np.testing.assert_allclose(d, -_bfgs_matrix(s_hist, y_hist, n) @ g, rtol=1e-10, atol=1e-12)
    # secant condition on the newest pair: H y = s
    np.testing.assert_allclose(lbfgs_direction(y_hist[-1], s_hist, y_hist, rho), -s_hist[-1], rtol=1e-10)
    assert np.array_equal(lbfgs_direction(g, [], [], []), -g)


class _LinearPhysics:
    """Stand-in for WavePhysics on the CPU: traces = band-scaled linear map of
    the field, misfit 0.5 ||p - d||^2 / ||d||^2 with its exact gradient."""

    def __init__(self, seed=0):
        g = torch.Generator().manual_seed(seed)
        self.A = torch.randn(40, 70 * 70, generator=g, dtype=torch.float64) / 70.0
        self.device = torch.device("cpu")

    def band_wavelet(self, band):
        return np.full(4, 1.0 if band is None else band / 20.0, np.float32)

    def bandlimit_to(self, traces, band):
        return traces * (1.0 if band is None else band / 20.0)

    def _pred(self, fields, wavelets):
        return (fields.double().reshape(fields.shape[0], -1) @ self.A.T) * wavelets[:, :1].double()

    def simulate_torch(self, fields, wavelets):
        return self._pred(fields, wavelets).float()

    def loss_and_grad_torch(self, fields, obs, wavelets):
        p = self._pred(fields, wavelets)
        d = p - obs.double().reshape(p.shape)
        e = (obs.double().reshape(p.shape) ** 2).sum(1)
        loss = 0.5 * (d * d).sum(1) / e
        grad = ((d / e[:, None]) * wavelets[:, :1].double()) @ self.A
        return loss.float(), grad.reshape(fields.shape).float()


def _linear_problem(phys, seed):
    gen = torch.Generator().manual_seed(seed)
    base = np.full((70, 70), 2.5, np.float32)
    tp = random_scene(gen, 6)
    true = torch.from_numpy(base) + R.render_residual(tp)[0]
    obs = phys.simulate_torch(true[None], torch.ones(1, 4)).numpy()[0]
    init = tp.clone()
    init[..., 5] *= 0.5
    return base, init[0].numpy(), obs


def test_refine_batch_driver_on_a_linear_problem():
    """Lockstep driver == one model at a time; the budget holds; the
    preconditioned arm falls back to plain gradients when its step fails."""
    from wgnet_amd.refine import refine_batch

    phys = _LinearPhysics()
    probs = [_linear_problem(phys, s) for s in (1, 2, 3)]
    dense = np.stack([p[0] for p in probs])
    init = np.stack([p[1] for p in probs])
    obs = np.stack([p[2] for p in probs])
    bands = (8.0, None)
    batch = refine_batch(dense, init, obs, phys, budget_solves=30, bands_hz=bands)
    for k in range(3):
        one = refine_batch(dense[k:k + 1], init[k:k + 1], obs[k:k + 1], phys, budget_solves=30, bands_hz=bands)[0]
        assert batch[k].solves == one.solves <= 32
        assert batch[k].final_loss == one.final_loss
        assert batch[k].final_loss < batch[k].initial_loss
    # a "preconditioner" that points uphill.  replace mode: every
    # preconditioned line search fails, so each model switches to plain
    # gradients and still descends.  first mode: the uphill direction is
    # never taken, so the run is plain L-BFGS's exactly.
    anti = refine_batch(dense, init, obs, phys, budget_solves=30, bands_hz=bands,
                        precondition=lambda g, f: -g, pc_mode="replace")
    for k in range(3):
        assert not anti[k].preconditioned
        assert any(h["preconditioned"] and not h["accepted"] for h in anti[k].history)
        assert anti[k].final_loss < anti[k].initial_loss
    first = refine_batch(dense, init, obs, phys, budget_solves=30, bands_hz=bands,
                         precondition=lambda g, f: -g, pc_mode="first")
    for k in range(3):
        assert first[k].final_loss == batch[k].final_loss and first[k].solves == batch[k].solves


def test_fit_update_recovers_a_representable_update():
    """The least-squares fit of a velocity update onto the scene parameters
    reproduces an update that the scene can represent exactly (a step in
    the coarse background and the Gaussian amplitudes, where the scene is
    linear)."""
    from wgnet_amd.refine import _fit_update, _Group, _Request, _Scene

    gen = torch.Generator().manual_seed(5)
    tp = random_scene(gen, 6)[0].numpy()
    active = np.flatnonzero(np.abs(tp[:, 5]) >= R.ACTIVE_AMP_KM)
    grp = _Group(np.full((1, 70, 70), 2.5, np.float32), active, "cpu")
    sc = _Scene(grp, 0, tp)
    x = sc.project(sc.x0)
    rng = np.random.default_rng(0)
    u = np.zeros_like(x)
    u[:81] = rng.standard_normal(81) * 0.05
    u[81 + 5::6] = rng.standard_normal(len(active)) * 0.02
    with torch.no_grad():
        want = grp.field_of([0], torch.tensor((x + u)[None], dtype=torch.float32)) - \
            grp.field_of([0], torch.tensor(x[None], dtype=torch.float32))
    pending = {0: _Request(True, x, None, False, True)}
    got_u = _fit_update([sc], pending, [0], want, iters=200)[0]
    with torch.no_grad():
        got = grp.field_of([0], torch.tensor((x + got_u)[None], dtype=torch.float32)) - \
            grp.field_of([0], torch.tensor(x[None], dtype=torch.float32))
    rel = float((got - want).norm() / want.norm())
    assert rel <= 1e-3, rel


def test_jump_steps_on_a_linear_problem():
    """An exact update hook (true - current field) makes the first step of a
    band land near the truth: lower misfit than plain L-BFGS at the same
    budget, with jump steps in the history; an update pointing uphill is
    never taken, so that run is plain L-BFGS's exactly."""
    from wgnet_amd.refine import refine_batch

    phys = _LinearPhysics()
    probs, trues = [], []
    for s in (1, 2, 3):
        gen = torch.Generator().manual_seed(s)
        base = np.full((70, 70), 2.5, np.float32)
        tp = random_scene(gen, 6)
        trues.append(torch.from_numpy(base) + R.render_residual(tp)[0])
        probs.append(_linear_problem(phys, s))
    dense = np.stack([p[0] for p in probs])
    init = np.stack([p[1] for p in probs])
    obs = np.stack([p[2] for p in probs])
    true = torch.stack(trues)
    bands = (8.0, None)
    plain = refine_batch(dense, init, obs, phys, budget_solves=20, bands_hz=bands)

    def exact(g, f, ids):
        return true[torch.as_tensor(ids)] - f

    jump = refine_batch(dense, init, obs, phys, budget_solves=20, bands_hz=bands, update=exact)
    for k in range(3):
        assert any(h["jump"] and h["accepted"] for h in jump[k].history)
        assert jump[k].final_loss < plain[k].final_loss
        assert jump[k].solves <= 22
    uphill = refine_batch(dense, init, obs, phys, budget_solves=20, bands_hz=bands,
                          update=lambda g, f, ids: f - true[torch.as_tensor(ids)])
    for k in range(3):
        assert uphill[k].final_loss == plain[k].final_loss and uphill[k].solves == plain[k].solves
