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
