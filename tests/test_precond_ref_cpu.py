"""The CPU reference of the preconditioner network pinned against torch
autograd in float64 (no bf16 rounding) and against a brute-force convolution."""

from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

from oracle import precond_ref as R


def _rand_params(rng, scale=0.3):
    n = sum(co * ci * 9 + co for ci, co in R.LAYERS)
    return (rng.standard_normal(n) * scale).astype(np.float64)


def test_backward_equals_autograd_float64():
    rng = np.random.default_rng(0)
    p = _rand_params(rng)
    x = rng.standard_normal((2, 2, 9, 11))
    t = rng.standard_normal((2, 9, 11))
    loss, g = R.loss_and_grad(p, x, t, bf16=False)
    layers = [(w.clone().requires_grad_(True), b.clone().requires_grad_(True)) for w, b in R.unflatten(p)]
    a = torch.from_numpy(x)
    for i, (w, b) in enumerate(layers):
        a = F.conv2d(a, w, b, padding=1)
        if i < 3:
            a = F.gelu(a)
    lt = ((a[:, 0] - torch.from_numpy(t)) ** 2).mean()
    lt.backward()
    ref = torch.cat([torch.cat([w.grad.reshape(-1), b.grad.reshape(-1)]) for w, b in layers]).numpy()
    assert abs(loss - float(lt)) <= 1e-12 * abs(float(lt))
    assert np.abs(g - ref).max() <= 1e-10 * np.abs(ref).max()


def test_first_layer_equals_brute_force_correlation():
    """conv2d with padding 1 is the cross-correlation the kernels compute:
    out[co, y, x] = b[co] + sum W[co, ci, ky, kx] in[ci, y + ky - 1, x + kx - 1]."""
    rng = np.random.default_rng(1)
    w = rng.standard_normal((3, 2, 3, 3))
    b = rng.standard_normal(3)
    x = rng.standard_normal((1, 2, 5, 6))
    out = F.conv2d(torch.from_numpy(x), torch.from_numpy(w), torch.from_numpy(b), padding=1).numpy()
    xp = np.pad(x, ((0, 0), (0, 0), (1, 1), (1, 1)))
    bf = np.zeros_like(out)
    for co in range(3):
        for yy in range(5):
            for xx in range(6):
                bf[0, co, yy, xx] = b[co] + np.sum(w[co] * xp[0, :, yy:yy + 3, xx:xx + 3])
    assert np.abs(out - bf).max() <= 1e-12


def test_bf16_rounding_is_nearest_even():
    vals = torch.tensor([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5], dtype=torch.float64)
    got = R.bf16_round(vals).numpy()
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3 * 2^-8 is a tie -> even (1 + 2^-6); 2^-9 rounds down
    assert got.tolist() == [1.0, 1.0 + 2 ** -6, 1.0, -2.5]
