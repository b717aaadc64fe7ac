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
    x = rng.standard_normal((2, 2, 21, 23))  # larger than the biggest dilation
    t = rng.standard_normal((2, 21, 23))
    loss, g = R.loss_and_grad(p, x, t, bf16=False)
    layers = [(w.clone().requires_grad_(True), b.clone().requires_grad_(True)) for w, b in R.unflatten(p)]
    a = torch.from_numpy(x)
    for i, (w, b) in enumerate(layers):
        d = R.DILATIONS[i]
        a = F.conv2d(a, w, b, padding=d, dilation=d)
        if i < 3:
            a = F.gelu(a)
    lt = ((a[:, 0] - torch.from_numpy(t)) ** 2).mean()
    lt.backward()
    ref = torch.cat([torch.cat([w.grad.reshape(-1), b.grad.reshape(-1)]) for w, b in layers]).numpy()
    assert abs(loss - float(lt)) <= 1e-12 * abs(float(lt))
    assert np.abs(g - ref).max() <= 1e-10 * np.abs(ref).max()


def test_dilated_layer_equals_brute_force_correlation():
    """conv2d with dilation D and padding D is the cross-correlation the
    kernels compute: out[co, y, x] = b[co] + sum W[co, ci, ky, kx]
    in[ci, y + (ky - 1) D, x + (kx - 1) D], zero outside the image."""
    rng = np.random.default_rng(1)
    w = rng.standard_normal((3, 2, 3, 3))
    b = rng.standard_normal(3)
    x = rng.standard_normal((1, 2, 7, 9))
    for D in (1, 2, 4):
        out = F.conv2d(torch.from_numpy(x), torch.from_numpy(w), torch.from_numpy(b), padding=D, dilation=D).numpy()
        bf = np.zeros_like(out)
        for co in range(3):
            for yy in range(7):
                for xx in range(9):
                    acc = b[co]
                    for ky in range(3):
                        for kx in range(3):
                            sy, sx = yy + (ky - 1) * D, xx + (kx - 1) * D
                            if 0 <= sy < 7 and 0 <= sx < 9:
                                acc += np.dot(w[co, :, ky, kx], x[0, :, sy, sx])
                    bf[0, co, yy, xx] = acc
        assert np.abs(out - bf).max() <= 1e-12, D


def test_bf16_rounding_is_nearest_even():
    vals = torch.tensor([1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5], dtype=torch.float64)
    got = R.bf16_round(vals).numpy()
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3 * 2^-8 is a tie -> even (1 + 2^-6); 2^-9 rounds down
    assert got.tolist() == [1.0, 1.0 + 2 ** -6, 1.0, -2.5]


def test_library_layer_table_equals_reference():
    """The HIP network's layers (host-side table, no GPU needed) are the ones
    the CPU reference restates."""
    import ctypes

    from wgnet_amd import _lib

    lib = _lib.load()
    assert lib.wg_pc_num_layers() == len(R.LAYERS)
    for l, ((ci, co), d) in enumerate(zip(R.LAYERS, R.DILATIONS)):
        v = [ctypes.c_int32() for _ in range(3)]
        assert lib.wg_pc_layer_shape(l, *[ctypes.byref(c) for c in v]) == 0
        assert [c.value for c in v] == [ci, co, d]
    assert lib.wg_pc_layer_shape(len(R.LAYERS), *[ctypes.byref(ctypes.c_int32()) for _ in range(3)]) != 0
    assert lib.wg_pc_num_params() == sum(co * ci * 9 + co for ci, co in R.LAYERS)
