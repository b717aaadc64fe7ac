"""The learned preconditioner's HIP kernels (MFMA bf16 forward, data and
weight gradients, AdamW) against the float64 CPU reference that rounds to
bf16 at the same points.  Tolerances are stated per check: the remaining
differences are float32 accumulation order and the occasional bf16 rounding
flip that it causes."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _layer_slices():
    from oracle.precond_ref import LAYERS

    out, k = [], 0
    for ci, co in LAYERS:
        out.append(slice(k, k + co * ci * 9))
        k += co * ci * 9
        out.append(slice(k, k + co))
        k += co
    return out


@pytest.fixture(scope="module")
def net():
    from wgnet_amd.precond import Preconditioner

    return Preconditioner(max_batch=8, seed=3)


def _inputs(seed, n=3):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, 2, 70, 70)).astype(np.float32)
    t = rng.standard_normal((n, 70, 70)).astype(np.float32) * 0.5
    return x, t


def test_forward_matches_cpu_reference(net):
    from oracle import precond_ref as R

    rng = np.random.default_rng(11)
    p = (rng.standard_normal(net.num_params) * 0.2).astype(np.float32)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    x, _ = _inputs(1)
    y = net.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    ref, _ = R.forward(p, x)
    err = relerr(y, ref.numpy())
    print(f"precond forward: rel-L2 {err:.3e}")
    assert err <= 1e-3


def test_parameter_gradient_matches_cpu_reference(net):
    from oracle import precond_ref as R

    rng = np.random.default_rng(12)
    p = (rng.standard_normal(net.num_params) * 0.2).astype(np.float32)
    net.set_params(p)
    x, t = _inputs(2)
    loss, g = net.loss_grad(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    rl, rg = R.loss_and_grad(p, x, t)
    print(f"precond loss {loss:.6e} vs {rl:.6e}")
    assert abs(loss - rl) <= 1e-4 * abs(rl)
    for k, sl in enumerate(_layer_slices()):
        err = relerr(g[sl], rg[sl])
        print(f"  layer {k // 2 + 1} {'W' if k % 2 == 0 else 'b'}: rel-L2 {err:.3e}")
        assert err <= 2e-3


def test_adamw_step_matches_formula(net):
    """First AdamW step: m^ = g, v^ = g^2, so theta1 = theta0 - lr (g / (|g| + eps) + wd theta0)."""
    rng = np.random.default_rng(13)
    p = (rng.standard_normal(net.num_params) * 0.2).astype(np.float32)
    net.set_params(p)
    x, t = _inputs(3)
    xd, td = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    _, g = net.loss_grad(xd, td)
    lr, wd = 1e-3, 1e-2
    net.train_step(xd, td, lr, wd)
    torch.cuda.synchronize()
    got = net.get_params().astype(np.float64)
    g64, p64 = g.astype(np.float64), p.astype(np.float64)
    want = p64 - lr * (g64 / (np.abs(g64) + 1e-8) + wd * p64)
    assert np.abs(got - want).max() <= 1e-6


def test_training_reduces_loss_on_a_smoothing_task(net):
    """Learn y = -(3x3 box mean of the first input channel): the loss falls
    by 10x in 300 steps."""
    rng = np.random.default_rng(14)
    net.set_params((rng.standard_normal(net.num_params) * 0.05).astype(np.float32))
    x = torch.from_numpy(rng.standard_normal((64, 2, 70, 70)).astype(np.float32)).cuda()
    box = torch.nn.functional.avg_pool2d(x[:, :1], 3, stride=1, padding=1, count_include_pad=True)[:, 0]
    target = -box
    first = None
    for k in range(300):
        idx = torch.from_numpy(rng.integers(0, 64, 8)).cuda()
        loss = float(net.train_step(x[idx], target[idx], 3e-3, 0.0))
        first = loss if first is None else first
    final, _ = net.loss_grad(x[:8], target[:8])
    print(f"smoothing task: loss {first:.4e} -> {final:.4e}")
    assert final < 0.1 * first
