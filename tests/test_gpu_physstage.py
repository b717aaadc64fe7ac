"""The phys-stage term on the GPU against the reference's per-model loop.

The reference (train.py:256-268 over physics.py:59-76) simulates the
observed data and evaluates the adjoint one model at a time through host
arrays; restated here with the same engine's one-model host calls
(`engine.simulate`, `engine.evaluate_adjoint`), then the same float32
balancing on the device.  The batched term must give the same loss, the
same scale and the same network gradient bit for bit, in every band of the
6, 9, 12 Hz, full schedule; the misfit mean within 1e-6 (numpy's float32
mean against torch's, a logged figure).  The first model's misfit and
gradient in the 6 Hz band are also checked against the CPU oracle."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _reference_term(ref, loss, field_km, true_km, nb, ratio, device):
    """train.py:260-268 with physics.py:59-76's per-model loops."""
    field = field_km[:nb]
    true_np = true_km[:nb].cpu().numpy()
    observed = np.empty((nb, ref.shots, ref.nt, 70), np.float32)
    for b in range(nb):
        observed[b] = ref.engine.simulate(np.ascontiguousarray(true_np[b].reshape(-1) * 1000.0, dtype=np.float32))
    f_np = field.detach().cpu().numpy()
    losses = np.zeros(nb, np.float32)
    grads = np.zeros((nb, 70, 70), np.float32)
    for b in range(nb):
        vel = np.ascontiguousarray(f_np[b].reshape(-1) * 1000.0, dtype=np.float32)
        obs = np.ascontiguousarray(observed[b].reshape(-1), dtype=np.float32)
        losses[b], g = ref.engine.evaluate_adjoint(vel, obs)
        grads[b] = g * 1000.0
    (g_sup,) = torch.autograd.grad(loss, field_km, retain_graph=True)
    g_phys = torch.from_numpy(grads).to(device) / nb
    scale = ratio * g_sup[:nb].norm() / (g_phys.norm() + 1e-12)
    return loss + scale * (field * g_phys).sum(), float(losses.mean()), float(scale), losses, grads, observed


def _network(seed, B, device):
    from wgnet_amd import synthetic

    torch.manual_seed(seed)
    w = (0.3 * torch.randn(B, 70, 70, device=device)).requires_grad_(True)
    start, true = synthetic.batch(B, seed=seed)
    base = torch.from_numpy(start / 1000.0).to(device)
    return w, base, torch.from_numpy((true / 1000.0).astype(np.float32)).to(device)


def test_phys_stage_equals_per_model_reference_loop():
    from wgnet_amd.physics import WavePhysics
    from wgnet_amd.physstage import PhysStage, band_at

    dev = torch.device("cuda")
    B, nb, steps = 5, 3, 4
    stage = PhysStage(steps=steps, bands="6,9,12,full", phys_ratio=0.5, phys_batch=nb)
    ref = WavePhysics()
    seen = []
    for step in range(1, steps + 1):
        w, base, true = _network(100 + step, B, dev)
        field = base + 0.2 * torch.tanh(w)
        loss = (field - true).abs().mean()
        new, parts = stage(loss, field, true, step)
        (gw,) = torch.autograd.grad(new, w)
        band = parts["band_hz"]
        seen.append(band)
        ref.set_band(band_at(step, steps, stage.bands))
        field_r = base + 0.2 * torch.tanh(w)
        loss_r = (field_r - true).abs().mean()
        new_r, misfit_r, scale_r, losses_r, grads_r, obs_r = _reference_term(ref, loss_r, field_r, true, nb, 0.5, dev)
        (gw_r,) = torch.autograd.grad(new_r, w)
        print(f"step {step} band {band}: loss {float(new.detach()):.9g} scale {float(parts['phys_scale']):.6g}",
              flush=True)
        assert float(new.detach()) == float(new_r.detach())
        assert float(parts["phys_scale"]) == scale_r
        assert torch.equal(gw, gw_r)
        assert abs(float(parts["phys_misfit"]) - misfit_r) <= 1e-6 * abs(misfit_r)
        if step == 1:  # 6 Hz: the first model against the CPU oracle
            from oracle import oracle as O

            g = O.Geometry(nt=2001, dt=0.0005)
            wav = ref.engine.wavelet()
            vel = np.ascontiguousarray(field_r[0].detach().cpu().numpy() * 1000.0, dtype=np.float32)
            tvel = np.ascontiguousarray(np.repeat(true[:1].cpu().numpy().reshape(1, -1) * 1000.0, 5, 0),
                                        dtype=np.float32)
            src = ref.engine.sources()
            obs_cpu = O.batch_forward_f32(g, src, tvel, wav, 5)
            assert obs_cpu.tobytes() == obs_r[0].tobytes()
            rl, rg = O.evaluate_adjoint_f32(g, src, vel, wav, obs_cpu)
            assert rl == float(losses_r[0])
            assert (rg * np.float32(1000.0)).tobytes() == grads_r[0].tobytes()
    assert seen == [6.0, 9.0, 12.0, 15.0]


@pytest.mark.parametrize("B", [1, 8, 51, 64])
def test_fused_observed_call_equals_two_calls(B):
    """wg_loss_grad_true_dev = wg_simulate_dev then wg_loss_grad_dev, bit for
    bit: beside each other (B <= 51 models, one round of units) and one after
    the other (64 models)."""
    from wgnet_amd import synthetic
    from wgnet_amd.physics import WavePhysics

    ph = WavePhysics(nt=600)
    ph.set_band(9.0)
    start, true = synthetic.batch(B, seed=900 + B)
    f = torch.from_numpy(start / 1000.0).cuda()
    t = torch.from_numpy((true / 1000.0).astype(np.float32)).cuda()
    l1, g1, o1 = ph.loss_and_grad_true_torch(f, t)
    o2 = ph.simulate_torch(t)
    l2, g2 = ph.loss_and_grad_torch(f, o2)
    torch.cuda.synchronize()
    assert torch.equal(o1, o2)
    assert torch.equal(l1, l2)
    assert torch.equal(g1, g2)
    # (0.3 s of recording: some anomalies lie deeper than the waves reach)
    assert bool(torch.isfinite(l1).all()) and float(l1.max()) > 0.0
