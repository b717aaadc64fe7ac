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


def test_per_model_wavelets_equal_engine_wavelet_calls():
    """wg_loss_grad_wavelets_dev / wg_simulate_wavelets_dev with each model's
    passband equal, bit for bit, calls made after set_band on the engine."""
    from wgnet_amd import synthetic
    from wgnet_amd.physics import WavePhysics

    ph = WavePhysics(shots=3, nt=500, dt=0.001)
    start, true = synthetic.batch(3, seed=77)
    f = torch.from_numpy(start / 1000.0).cuda()
    bands = [6.0, None, 12.0]
    obs_full = ph.simulate_torch(torch.from_numpy(true / 1000.0).cuda())
    obs = torch.stack([ph.bandlimit_to(obs_full[b:b + 1], bands[b])[0] for b in range(3)])
    wav = torch.stack([torch.from_numpy(ph.band_wavelet(bd)) for bd in bands]).cuda()
    loss_w, grad_w = ph.loss_and_grad_torch(f, obs, wavelets=wav)
    pred_w = ph.simulate_torch(f, wavelets=wav)
    for b, bd in enumerate(bands):
        ph.set_band(bd)
        loss_1, grad_1 = ph.loss_and_grad_torch(f[b:b + 1], obs[b:b + 1])
        pred_1 = ph.simulate_torch(f[b:b + 1])
        assert torch.equal(loss_w[b:b + 1], loss_1)
        assert torch.equal(grad_w[b:b + 1], grad_1)
        assert torch.equal(pred_w[b:b + 1], pred_1)
    ph.set_band(None)


def test_refine_batch_equals_single_refines():
    """The lockstep driver gives each model the run refine() gives it alone."""
    from test_refine_cpu import random_scene
    from wgnet_amd.physics import WavePhysics
    from wgnet_amd.refine import refine, refine_batch
    from wgnet_amd.render import render_residual

    physics = WavePhysics(shots=2, nt=601, dt=0.001)
    dense, init, obs = [], [], []
    for seed in (5, 6, 7):
        gen = torch.Generator().manual_seed(seed)
        base = np.full((70, 70), 2.0 + 0.2 * (seed - 5), np.float32)
        tp = random_scene(gen, 8)
        with torch.no_grad():
            obs.append(physics.simulate((torch.from_numpy(base) + render_residual(tp)).numpy())[0])
        ip = tp.clone()
        ip[..., 5] *= 0.5
        dense.append(base)
        init.append(ip[0].numpy())
    batch = refine_batch(np.stack(dense), np.stack(init), np.stack(obs), physics, budget_solves=24,
                         bands_hz=(9.0, None))
    for k in range(3):
        one = refine(dense[k], init[k], obs[k], physics, budget_solves=24, bands_hz=(9.0, None))
        print(f"model {k}: batch {batch[k].final_loss:.6e} ({batch[k].solves}), single {one.final_loss:.6e} ({one.solves})")
        assert batch[k].solves == one.solves
        assert [h["accepted"] for h in batch[k].history] == [h["accepted"] for h in one.history]
        assert abs(batch[k].final_loss - one.final_loss) <= 1e-9 * abs(one.final_loss)
        assert batch[k].final_loss < batch[k].initial_loss
