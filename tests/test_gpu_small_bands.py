"""Small batches of the 70 x 70 model as row bands: a call with at most 64
(model, shot) units runs each unit as 4 band workgroups, one with at most 128
units as 2, on band kernels compiled for the grid (the default) or on the
runtime-geometry band kernels; larger calls keep one workgroup per unit.

Checked, bit for bit against the whole-unit kernels (band split off): gathers,
losses and gradients at 1, 8 (4 bands), 13 (2 bands) and 60 models (whole
units), through the batched device path and the fused observed-data call
(whose two forwards run as bands on two streams at once); the host-pointer
path against the float32 oracle; and a wait that gives up (forced from the
host) turns every result of the call into NaN, with every kernel ending
normally."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("B", [1, 8, 13, 60])
def test_small_batch_bands_equal_whole_units(B):
    from wgnet_amd import synthetic
    from wgnet_amd.physics import WavePhysics

    ph = WavePhysics(nt=400, dt=0.001)
    lib, h = ph.engine._lib, ph.engine.handle
    start, true = synthetic.batch(B, seed=700 + B)
    f = torch.from_numpy(start / 1000.0).cuda()
    t = torch.from_numpy((true / 1000.0).astype(np.float32)).cuda()
    res = {}
    for mode in (1, 2, 0):  # compiled bands, runtime-geometry bands, whole units
        assert lib.wg_set_band_split(h, mode) == 0
        obs = ph.simulate_torch(t)
        l, g = ph.loss_and_grad_torch(f, obs)
        lf, gf, of = ph.loss_and_grad_true_torch(f, t)
        torch.cuda.synchronize()
        res[mode] = (obs, l, g, lf, gf, of)
    lib.wg_set_band_split(h, 1)
    obs, l, g = res[0][:3]
    assert bool(torch.isfinite(l).all()) and float(l.max()) > 0.0 and float(g.abs().max()) > 0.0
    for mode in (1, 2):
        for k, (a, b) in enumerate(zip(res[mode], res[0])):
            assert torch.equal(a, b), (B, mode, k)
    # the fused call equals the two calls
    assert torch.equal(res[1][5], obs) and torch.equal(res[1][3], l) and torch.equal(res[1][4], g)


def test_small_batch_bands_match_oracle(orc):
    """One model, 5 shots (4 bands) through the host-pointer API: gathers,
    loss and gradient equal the float32 oracle."""
    from wgnet_amd import synthetic
    from wgnet_amd.wave_gauss_rs import PyWaveEngine

    nt = 500
    E = PyWaveEngine(5, nt, 0.001, 15.0, 20, 1000.0, 5000.0)
    start, true = synthetic.batch(1, seed=4711)
    obs = E.simulate(true[0].reshape(-1))
    loss, grad = E.evaluate_adjoint(start[0].reshape(-1), obs.reshape(-1))
    for s, (x, z) in enumerate(E.sources()):
        g = orc.Geometry(nvx=70, nvz=70, nt=nt, dt=0.001, src_x=float(x), src_z=float(z))
        assert np.array_equal(obs[s], orc.forward_f32(g, true[0], E.wavelet())), s
    rl, rg = orc.evaluate_adjoint_f32(orc.Geometry(nvx=70, nvz=70, nt=nt, dt=0.001), E.sources(), start[0],
                                      E.wavelet(), obs)
    print(f"70x70, 1 model x 5 shots as bands: loss {loss:.9e} vs {rl:.9e}, gradient rel-L2 {relerr(grad, rg):.2e}",
          flush=True)
    assert loss == rl and np.array_equal(grad, rg)


@pytest.mark.parametrize("mode", [1, 2])
def test_small_batch_band_wait_that_gives_up_gives_nan(mode):
    from wgnet_amd import synthetic
    from wgnet_amd.physics import WavePhysics

    ph = WavePhysics(shots=2, nt=60, dt=0.001)
    lib, h = ph.engine._lib, ph.engine.handle
    assert lib.wg_set_band_split(h, mode) == 0
    start, true = synthetic.batch(2, seed=6)
    t = torch.from_numpy((true / 1000.0).astype(np.float32)).cuda()
    f = torch.from_numpy(start / 1000.0).cuda()
    obs = ph.simulate_torch(t)
    l0, g0 = ph.loss_and_grad_torch(f, obs)
    assert lib.wg_set_queue_wait_limit(h, -1) == 0
    bad = ph.simulate_torch(t)
    l1, g1 = ph.loss_and_grad_torch(f, obs)
    torch.cuda.synchronize()
    assert bool(torch.isnan(bad).all())
    assert bool(torch.isnan(l1).all()) and bool(torch.isnan(g1).all())
    assert lib.wg_set_queue_wait_limit(h, 1 << 24) == 0
    l2, g2 = ph.loss_and_grad_torch(f, obs)
    assert torch.equal(l2, l0) and torch.equal(g2, g0)
