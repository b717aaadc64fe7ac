"""The wave-equation term of WGNet's `phys` fine-tuning stage, batched on the GPU.

Counterpart of the phys block of the training loop (python/wgnet/train.py:256-268)
and its passband schedule (train.py:195-196, 229-230, 346-348).  Per training step:
  band       = bands[min(len(bands) - 1, int(len(bands) * (step - 1) / max(1, steps)))]
  observed   = the first nb true models simulated in that band
  losses, g  = misfit and gradient (w.r.t. km/s) of the first nb predicted fields
  g_phys     = g / nb
  scale      = phys_ratio * ||g_sup[:nb]|| / (||g_phys|| + 1e-12),  g_sup = d loss / d field
  loss      += scale * (field[:nb] * g_phys).sum()
so the physics gradient enters the network with phys_ratio times the norm of
the supervised gradient on the same fields.

What changes is where the solves run: the reference simulates the observed
data and evaluates the adjoint one model at a time through host arrays
(physics.py:59-76, a host copy of every gather and gradient); here the true
models and the predicted fields stay on the device, and both solves are one
batched engine call, the observed forward running beside the predicted one
when the batch fits one round of the device.  The arithmetic, including every
float32 operation of the balancing, is the reference's, so the loss and its
gradient are the same numbers.
"""

from __future__ import annotations

import torch

from .physics import WavePhysics

DEFAULT_BANDS = "6,9,12,full"  # train.py:348
DEFAULT_RATIO = 0.5  # train.py:346
DEFAULT_BATCH = 8  # train.py:347


def parse_bands(spec: str) -> list[float | None]:
    """train.py:195-196: "6,9,12,full" -> [6.0, 9.0, 12.0, None]."""
    return [None if b.strip() in ("full", "none") else float(b) for b in spec.split(",")]


def band_at(step: int, steps: int, bands: list[float | None]) -> float | None:
    """The passband of training step `step` (1-based) of `steps` (train.py:257)."""
    return bands[min(len(bands) - 1, int(len(bands) * (step - 1) / max(1, steps)))]


def physics_term(physics, loss: torch.Tensor, field_km: torch.Tensor, true_km: torch.Tensor, nb: int,
                 phys_ratio: float = DEFAULT_RATIO) -> tuple[torch.Tensor, dict]:
    """Adds the balanced wave-equation term to `loss` (train.py:260-268) in the
    physics' current band.  `field_km` is the network's composed field
    [B, 70, 70] (part of loss's graph), `true_km` the true models of the batch
    [>= nb, 70, 70] on the same device.  Returns (new loss, parts) with parts
    "phys_misfit" and "phys_scale" as 0-d tensors."""
    nb = min(int(nb), field_km.shape[0])
    if nb <= 0:
        raise ValueError("phys batch must be positive")
    field = field_km[:nb]
    with torch.no_grad():
        phys_losses, phys_grads, _ = physics.loss_and_grad_true_torch(field.detach(), true_km[:nb])
    (g_sup,) = torch.autograd.grad(loss, field_km, retain_graph=True)
    g_phys = phys_grads / nb
    scale = phys_ratio * g_sup[:nb].norm() / (g_phys.norm() + 1e-12)
    loss = loss + scale * (field * g_phys).sum()
    return loss, {"phys_misfit": phys_losses.mean().detach(), "phys_scale": scale.detach()}


class PhysStage:
    """The phys stage's physics, schedule and balancing: call once per
    training step with the step's supervised loss, the predicted field and
    the batch's true models (train.py:256-268)."""

    def __init__(self, steps: int, bands: str | list = DEFAULT_BANDS, phys_ratio: float = DEFAULT_RATIO,
                 phys_batch: int = DEFAULT_BATCH, physics: WavePhysics | None = None):
        self.steps = int(steps)
        self.bands = parse_bands(bands) if isinstance(bands, str) else list(bands)
        if not self.bands:
            raise ValueError("empty passband schedule")
        self.phys_ratio, self.phys_batch = float(phys_ratio), int(phys_batch)
        self.physics = physics if physics is not None else WavePhysics()

    def __call__(self, loss: torch.Tensor, field_km: torch.Tensor, true_km: torch.Tensor,
                 step: int) -> tuple[torch.Tensor, dict]:
        band = band_at(step, self.steps, self.bands)
        if self.physics.fmax != band:
            self.physics.set_band(band)
        loss, parts = physics_term(self.physics, loss, field_km, true_km, self.phys_batch, self.phys_ratio)
        parts["band_hz"] = band or 15.0
        return loss, parts


__all__ = ["DEFAULT_BANDS", "DEFAULT_RATIO", "DEFAULT_BATCH", "parse_bands", "band_at", "physics_term", "PhysStage"]
