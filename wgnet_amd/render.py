"""Additive local-window Gaussian scene used by `refine()`.

Counterpart of the parts of python/wgnet/render.py that refinement uses
(render.py:20-31 constants, :64-83 activation / clamp, :86-109 render).
A 70 x 70 grid carries one slot per 2 x 2 pixel cell (35 x 35 slots); slot
(iz, ix) is anchored at (x, z) = (2 ix + 1, 2 iz + 1) in pixel-edge coordinates
(pixel centres at k + 0.5) and owns six parameters
(dx, dz, log_sx, log_sz, angle, amp).  It adds
    amp * exp(-(u^2 + v^2) / 2),
    u = ((px - cx) cos a + (pz - cz) sin a) / sx,
    v = ((pz - cz) cos a - (px - cx) sin a) / sz,
to the pixels within 12 of its anchor pixel, with (cx, cz) = anchor + (dx, dz).
"""

from __future__ import annotations

import math

import torch

GRID = 70
SLOT_SIDE = 35
N_SLOTS = SLOT_SIDE * SLOT_SIDE
WINDOW = 12
OFFSET_MAX = 2.0
SIGMA_MIN = 0.5
SIGMA_MAX = 6.0
AMP_MAX = 1.0
ACTIVE_AMP_KM = 0.05
N_GAUSS_PARAMS = 6


def activate_gaussians(raw: torch.Tensor) -> torch.Tensor:
    """Unconstrained [..., 6] -> bounded physical parameters (render.py:64-73)."""
    lo, hi = math.log(SIGMA_MIN), math.log(SIGMA_MAX)
    return torch.stack(
        [
            OFFSET_MAX * torch.tanh(raw[..., 0]),
            OFFSET_MAX * torch.tanh(raw[..., 1]),
            lo + (hi - lo) * torch.sigmoid(raw[..., 2]),
            lo + (hi - lo) * torch.sigmoid(raw[..., 3]),
            0.5 * math.pi * torch.tanh(raw[..., 4]),
            AMP_MAX * torch.tanh(raw[..., 5]),
        ],
        dim=-1,
    )


def clamp_gaussians(params: torch.Tensor) -> torch.Tensor:
    """Projection onto the parameter bounds (render.py:76-83)."""
    lo = torch.tensor([-OFFSET_MAX, -OFFSET_MAX, math.log(SIGMA_MIN), math.log(SIGMA_MIN), -0.5 * math.pi, -AMP_MAX],
                      dtype=params.dtype, device=params.device)
    hi = torch.tensor([OFFSET_MAX, OFFSET_MAX, math.log(SIGMA_MAX), math.log(SIGMA_MAX), 0.5 * math.pi, AMP_MAX],
                      dtype=params.dtype, device=params.device)
    return torch.clamp(params, min=lo, max=hi)


_GEOM: dict = {}


def _slot_geometry(slots: torch.Tensor):
    """Per slot: flat pixel index [n, W^2] (GRID*GRID for pixels off the grid),
    pixel centres [n, W^2, 2] and validity [n, W^2]."""
    key = (str(slots.device), tuple(slots.tolist()))
    hit = _GEOM.get(key)
    if hit is not None:
        return hit
    iz, ix = slots // SLOT_SIDE, slots % SLOT_SIDE
    off = torch.arange(-WINDOW, WINDOW + 1, device=slots.device)
    pz = (2 * iz)[:, None, None] + off[None, :, None]
    px = (2 * ix)[:, None, None] + off[None, None, :]
    pz, px = pz.expand(-1, off.numel(), off.numel()).reshape(len(slots), -1), px.expand(-1, off.numel(), off.numel()).reshape(len(slots), -1)
    valid = (pz >= 0) & (pz < GRID) & (px >= 0) & (px < GRID)
    flat = torch.where(valid, pz * GRID + px, torch.full_like(pz, GRID * GRID))
    centres = torch.stack([px.float() + 0.5, pz.float() + 0.5], dim=-1)
    anchors = torch.stack([2.0 * ix.float() + 1.0, 2.0 * iz.float() + 1.0], dim=-1)
    out = (flat, centres, valid, anchors)
    _GEOM[key] = out
    return out


def render_residual(params: torch.Tensor, slots: torch.Tensor | None = None) -> torch.Tensor:
    """[B, n, 6] activated parameters of `slots` (default: all 1225 in order)
    -> [B, 70, 70] km/s residual field (render.py:86-109)."""
    bsz, n = params.shape[0], params.shape[1]
    if slots is None:
        slots = torch.arange(N_SLOTS, device=params.device)
    flat, centres, valid, anchors = _slot_geometry(slots.to(params.device))
    c = anchors.to(params.dtype)[None] + params[..., 0:2]  # [B, n, 2]
    d = centres.to(params.dtype)[None] - c[:, :, None, :]  # [B, n, W^2, 2]
    cos_a, sin_a = torch.cos(params[..., 4])[..., None], torch.sin(params[..., 4])[..., None]
    u = (d[..., 0] * cos_a + d[..., 1] * sin_a) * torch.exp(-params[..., 2])[..., None]
    v = (d[..., 1] * cos_a - d[..., 0] * sin_a) * torch.exp(-params[..., 3])[..., None]
    val = params[..., 5][..., None] * torch.exp(-0.5 * (u * u + v * v)) * valid.to(params.dtype)[None]
    field = torch.zeros(bsz, GRID * GRID + 1, dtype=params.dtype, device=params.device)
    field.index_add_(1, flat.reshape(-1), val.reshape(bsz, -1))
    return field[:, : GRID * GRID].view(bsz, GRID, GRID)
