"""Row bands: the 100 x 100 and 141 x 141 models (144 x 144 and 185 x 185
padded), too large for one workgroup, run as 2 and 3 row bands, one workgroup
each, exchanging edge rows every step, on band kernels compiled for these
grids (the default) or on the runtime-geometry band kernels.  A 141 x 141
call runs 4 bands a unit instead where that fills the compute units in fewer
rounds (a call of few units).

Checked: gathers, loss and gradient against the float32 oracle (gathers rel-L2
<= 1e-5, loss |rel| <= 1e-6, gradient rel-L2 <= 1e-6; the per-cell arithmetic
is the whole-grid kernels', so bit-equality is expected), and bit for bit
between the compiled band kernels, the runtime-geometry band kernels and the
HBM-streamed kernels on the same engine; off-node sources in several bands
(the 141 x 141 call of 3 units runs 4 bands a unit); a batch with more bands
than compute units (150 units of 3 bands); and a wait that gives up (forced
from the host) turns every result of the call into NaN, on both band
kernels."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("n,bands,nt", [(100, 2, 400), (141, 3, 400)])
def test_band_kernels_match_oracle_and_streamed(orc, n, bands, nt):
    from wgnet_amd import synthetic
    from wgnet_amd.wave_gauss_rs import PyWaveEngine

    S = 3
    E = PyWaveEngine(S, nt, 0.001, 15.0, 20, 1000.0, 5000.0, nvx=n, nvz=n)
    lib, h = E._lib, E.handle
    assert E.streamed and lib.wg_engine_bands(h) == bands
    # sources in the top band, one off-node, one deeper (in a lower band)
    src = np.array([[55.0, 10.0], [0.5 * (n - 1) * 10.0 + 3.3, 17.5], [(n - 1) * 10.0 - 40.0, 0.62 * n * 10.0]],
                   np.float32)
    E.set_sources(src)
    _, true = synthetic.batch(1, seed=910 + n, nvz=n, nvx=n)
    true = true[0]
    start = true.copy()
    start[8:24, n // 3:n // 3 + 25] += 130.0  # shallow: a misfit the gathers see within nt
    obs = E.simulate(true.reshape(-1))
    loss, grad = E.evaluate_adjoint(start.reshape(-1), obs.reshape(-1))
    assert lib.wg_set_band_split(h, 2) == 0  # runtime-geometry band kernels
    obs_r = E.simulate(true.reshape(-1))
    loss_r, grad_r = E.evaluate_adjoint(start.reshape(-1), obs.reshape(-1))
    assert lib.wg_set_band_split(h, 0) == 0  # HBM-streamed kernels
    obs_s = E.simulate(true.reshape(-1))
    loss_s, grad_s = E.evaluate_adjoint(start.reshape(-1), obs.reshape(-1))
    assert lib.wg_set_band_split(h, 1) == 0
    print(f"{n}x{n}, {bands} bands: gathers equal to runtime bands {np.array_equal(obs, obs_r)} and streamed "
          f"{np.array_equal(obs, obs_s)}, loss {loss == loss_r} {loss == loss_s}, gradient "
          f"{np.array_equal(grad, grad_r)} {np.array_equal(grad, grad_s)}", flush=True)
    for s, (x, z) in enumerate(src):
        g = orc.Geometry(nvx=n, nvz=n, nt=nt, dt=0.001, src_x=float(x), src_z=float(z))
        err = relerr(obs[s], orc.forward_f32(g, true, E.wavelet()))
        print(f"  shot {s} at ({x}, {z}): gathers rel-L2 {err:.2e} vs FP32 oracle", flush=True)
        assert err <= 1e-5
    g = orc.Geometry(nvx=n, nvz=n, nt=nt, dt=0.001)
    rl, rg = orc.evaluate_adjoint_f32(g, src, start, E.wavelet(), obs)
    print(f"  loss {loss:.9e} vs {rl:.9e}; gradient rel-L2 {relerr(grad, rg):.2e}", flush=True)
    assert abs(loss - rl) <= 1e-6 * rl
    assert relerr(grad, rg) <= 1e-6
    assert np.array_equal(obs, obs_r) and loss == loss_r and np.array_equal(grad, grad_r)
    assert np.array_equal(obs, obs_s) and loss == loss_s and np.array_equal(grad, grad_s)


def test_band_batch_larger_than_the_device():
    """More bands than compute units (workgroups start as others finish),
    batched device path: equal to the streamed kernels bit for bit."""
    from wgnet_amd import synthetic
    from wgnet_amd.physics import WavePhysics

    n, nt, B, S = 141, 300, 30, 5  # 150 units x 3 bands = 450 workgroups (two rounds; 4 bands would take three)
    ph = WavePhysics(shots=S, nt=nt, dt=0.001, nvx=n, nvz=n)
    lib, h = ph.engine._lib, ph.engine.handle
    assert lib.wg_engine_bands(h) == 3
    start, true = synthetic.batch(B, seed=77, nvz=n, nvx=n)
    t_km = torch.from_numpy(true / 1000.0).cuda()
    f_km = torch.from_numpy(start / 1000.0).cuda()
    obs = ph.simulate_torch(t_km)
    l, g = ph.loss_and_grad_torch(f_km, obs)
    results = {}
    for mode in (2, 0):  # runtime-geometry band kernels, HBM-streamed kernels
        assert lib.wg_set_band_split(h, mode) == 0
        o = ph.simulate_torch(t_km)
        results[mode] = (o,) + tuple(ph.loss_and_grad_torch(f_km, o))
    lib.wg_set_band_split(h, 1)
    torch.cuda.synchronize()
    assert bool(torch.isfinite(l).all()) and bool((l >= 0).all()) and float(l.max()) > 0.0
    for mode, (o, l_m, g_m) in results.items():
        assert torch.equal(obs, o) and torch.equal(l, l_m) and torch.equal(g, g_m), mode


@pytest.mark.parametrize("mode", [1, 2])
def test_band_wait_that_gives_up_gives_nan(mode):
    """A negative wait limit makes every band's first exchange give up (all
    kernels end normally): the gathers, losses and gradients of the call are
    NaN; the default limit gives the results back."""
    from wgnet_amd import synthetic
    from wgnet_amd.physics import WavePhysics

    n, nt = 100, 60
    ph = WavePhysics(shots=2, nt=nt, dt=0.001, nvx=n, nvz=n)
    lib, h = ph.engine._lib, ph.engine.handle
    assert lib.wg_set_band_split(h, mode) == 0
    start, true = synthetic.batch(2, seed=5, nvz=n, nvx=n)
    obs = ph.simulate_torch(torch.from_numpy(true / 1000.0).cuda())
    f = torch.from_numpy(start / 1000.0).cuda()
    l0, g0 = ph.loss_and_grad_torch(f, obs)
    assert lib.wg_set_queue_wait_limit(h, -1) == 0
    bad_obs = ph.simulate_torch(torch.from_numpy(true / 1000.0).cuda())
    l1, g1 = ph.loss_and_grad_torch(f, obs)
    torch.cuda.synchronize()
    assert bool(torch.isnan(bad_obs).all())
    assert bool(torch.isnan(l1).all()) and bool(torch.isnan(g1).all())
    assert lib.wg_set_queue_wait_limit(h, 1 << 24) == 0
    l2, g2 = ph.loss_and_grad_torch(f, obs)
    assert torch.equal(l2, l0) and torch.equal(g2, g0)
