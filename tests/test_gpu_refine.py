"""refine() on the GPU engine: the reference's own refinement check
(python/verify.py:107-117) and the budget / monotonicity invariants."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_refine_reduces_misfit_verify_problem():
    from test_refine_cpu import random_scene
    from wgnet_amd.physics import WavePhysics
    from wgnet_amd.refine import refine
    from wgnet_amd.render import render_residual

    physics = WavePhysics(shots=2, nt=801)  # data.py DT = 0.5 ms
    gen = torch.Generator().manual_seed(4)
    base = np.full((70, 70), 2.5, np.float32)
    true_params = random_scene(gen, 10)
    with torch.no_grad():
        observed = physics.simulate((torch.from_numpy(base) + render_residual(true_params)).numpy())[0]
    init = true_params.clone()
    init[..., 5] *= 0.5
    res = refine(base, init[0].numpy(), observed, physics, budget_solves=30, bands_hz=(10.0, None))
    print(f"refine: {res.initial_loss:.4e} -> {res.final_loss:.4e} in {res.solves} solves, {res.seconds:.2f}s")
    assert res.final_loss < 0.7 * res.initial_loss  # verify.py:117
    assert res.solves <= 30 + 2  # per-band budgets plus the two full-band misfit evaluations
    for band in (10.0, None):
        losses = [h["loss"] for h in res.history if h["band_hz"] == band]
        assert all(b <= a for a, b in zip(losses, losses[1:]))
