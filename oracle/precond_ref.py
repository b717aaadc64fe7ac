"""CPU reference of the learned preconditioner network (test infrastructure:
only tests/ may import it).

The network has no counterpart in the reference repository; it is the network
of dilated 3x3 convolutions 2 -> 32 -> 32 -> 32 -> 1 (dilations 4, 16, 8, 1,
GELU after the first three layers, zero padding equal to the dilation) of the
HIP kernels.  This module restates its forward
pass, mean-squared-error loss and parameter gradient in float64 with torch's
CPU convolution routines (torch.nn.functional.conv2d and torch.nn.grad.*),
rounding to bfloat16 at the same points as the HIP kernels (network input,
weights, activations, output gradient and pre-activation gradients), so that
the only remaining difference is float32 accumulation order on the GPU.

With `bf16=False` the backward pass is exactly the gradient of the float64
network; tests pin it against torch autograd.
"""

from __future__ import annotations

import numpy as np
import torch
import torch.nn.functional as F

LAYERS = ((2, 32), (32, 32), (32, 32), (32, 1))
DILATIONS = (4, 16, 8, 1)


def bf16_round(x: torch.Tensor) -> torch.Tensor:
    """Round-to-nearest-even to bfloat16, returned in x's dtype."""
    return x.to(torch.float32).to(torch.bfloat16).to(x.dtype)


def unflatten(params: np.ndarray) -> list[tuple[torch.Tensor, torch.Tensor]]:
    """Flat PyTorch-order vector -> [(W [co, ci, 3, 3], b [co])] in float64."""
    out, k = [], 0
    for ci, co in LAYERS:
        w = torch.from_numpy(np.asarray(params[k:k + co * ci * 9], np.float64).reshape(co, ci, 3, 3).copy())
        k += co * ci * 9
        b = torch.from_numpy(np.asarray(params[k:k + co], np.float64).copy())
        k += co
        out.append((w, b))
    assert k == len(params)
    return out


def _gelu(x):
    return 0.5 * x * (1.0 + torch.erf(x / np.sqrt(2.0)))


def _gelu_d(x):
    return 0.5 * (1.0 + torch.erf(x / np.sqrt(2.0))) + x * torch.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def forward(params: np.ndarray, x: np.ndarray, bf16: bool = True):
    """x [N, 2, H, W] -> (y [N, H, W], cache)."""
    r = bf16_round if bf16 else (lambda t: t)
    layers = unflatten(params)
    a = r(torch.from_numpy(np.asarray(x, np.float64)))
    acts, pres, ws = [a], [], []
    for li, (w, b) in enumerate(layers):
        wq = r(w)
        ws.append(wq)
        d = DILATIONS[li]
        pre = F.conv2d(a, wq, b, padding=d, dilation=d)
        pres.append(pre)
        if li < len(layers) - 1:
            a = r(_gelu(pre))
            acts.append(a)
    return pres[-1][:, 0], (acts, pres, ws)


def loss_and_grad(params: np.ndarray, x: np.ndarray, target: np.ndarray, bf16: bool = True):
    """(mean squared error, flat parameter gradient in PyTorch order)."""
    r = bf16_round if bf16 else (lambda t: t)
    y, (acts, pres, ws) = forward(params, x, bf16)
    t = torch.from_numpy(np.asarray(target, np.float64))
    P = y.numel()
    d = y - t
    loss = float((d * d).sum() / P)
    g = r((2.0 * d / P)[:, None])
    grads = []
    for li in range(len(LAYERS) - 1, -1, -1):
        a_in = acts[li]
        d = DILATIONS[li]
        gw = torch.nn.grad.conv2d_weight(a_in, ws[li].shape, g, padding=d, dilation=d)
        gb = g.sum((0, 2, 3))
        grads.append((gw, gb))
        if li > 0:
            gin = torch.nn.grad.conv2d_input(a_in.shape, ws[li], g, padding=d, dilation=d)
            g = r(gin * _gelu_d(pres[li - 1]))
    flat = []
    for gw, gb in reversed(grads):
        flat.append(gw.reshape(-1))
        flat.append(gb.reshape(-1))
    return loss, torch.cat(flat).numpy()
