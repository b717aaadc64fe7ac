"""Edge cases against the CPU oracle: NaN velocities (the reference's
f32::clamp keeps NaN, geology-core lib.rs:101-104), one- and few-step runs,
sources on and beyond the model edges, and empty batches.  Tolerances: the
kernels keep the reference's operation order, so gathers, losses and
gradients are compared for equality (NaN positions included)."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _layers():
    m = np.full((70, 70), 1900.0, np.float32)
    m[25:] = 2700.0
    m[48:] = 3600.0
    return m


@pytest.mark.parametrize("grid", [70, 100])
def test_nan_velocity_propagates_like_the_cpu_solver(orc, grid):
    """A NaN cell makes the same samples NaN as in the CPU solver (resident
    70x70 and streamed 100x100 kernels)."""
    from wgnet_amd.wave_gauss_rs import PyWaveEngine

    nt = 120
    E = PyWaveEngine(1, nt, 0.001, 15.0, 20, 1000.0, 5000.0, nvx=grid, nvz=grid)
    E.set_sources(np.array([[200.0, 10.0]], np.float32))
    vel = np.full((grid, grid), 2500.0, np.float32)
    vel[grid // 2, grid // 2] = np.nan
    gpu = E.simulate(vel.reshape(-1))[0]
    ref = orc.forward_f32(orc.Geometry(nvx=grid, nvz=grid, nt=nt, dt=0.001, src_x=200.0, src_z=10.0), vel,
                          E.wavelet())
    assert np.isnan(ref).any() and not np.isnan(ref).all()
    assert np.array_equal(gpu, ref, equal_nan=True)


@pytest.mark.parametrize("nt", [1, 2, 5])
def test_few_step_loss_and_gradient(orc, nt):
    """nt = 1, 2, 5: the time loops' first and last steps, with random observed
    data so that the energy is positive."""
    from wgnet_amd.wave_gauss_rs import PyWaveEngine

    E = PyWaveEngine(2, nt, 0.001, 15.0, 20, 1000.0, 5000.0)
    E.set_sources(np.array([[100.0, 10.0], [455.0, 15.0]], np.float32))
    rng = np.random.default_rng(nt)
    obs = rng.standard_normal(2 * nt * 70).astype(np.float32)
    vel = _layers()
    loss, grad = E.evaluate_adjoint(vel.reshape(-1), obs)
    rl, rg = orc.evaluate_adjoint_f32(orc.Geometry(nt=nt, dt=0.001), E.sources(), vel, E.wavelet(),
                                      obs.reshape(2, nt, 70))
    assert loss == rl
    assert np.array_equal(grad, rg)


def test_sources_on_and_beyond_the_edges(orc):
    """Sources at the model corners and up to 1.5 cells outside it (inside the
    C-PML pad), fractional positions included."""
    from wgnet_amd.wave_gauss_rs import PyWaveEngine

    src = np.array([[-15.0, 0.0], [0.0, 10.0], [345.5, 695.0], [690.0, 10.0], [705.0, 700.0]], np.float32)
    nt = 400
    E = PyWaveEngine(len(src), nt, 0.001, 15.0, 20, 1000.0, 5000.0)
    E.set_sources(src)
    vel = _layers()
    gpu = E.simulate(vel.reshape(-1))
    for s, (x, z) in enumerate(src):
        ref = orc.forward_f32(orc.Geometry(nt=nt, dt=0.001, src_x=float(x), src_z=float(z)), vel, E.wavelet())
        assert np.abs(ref).max() > 0.0
        assert np.array_equal(gpu[s], ref), f"source {s} at ({x}, {z})"


@pytest.mark.parametrize("bands,fwd_block", [(True, 1), (False, 1), (False, 2)])
def test_single_step_streamed(orc, bands, fwd_block):
    """nt = 1, 2 and 3 on the 100 x 100 grid: row bands, or the HBM-streamed
    kernels with one or two forward steps per launch."""
    from wgnet_amd.wave_gauss_rs import PyWaveEngine

    for nt in (1, 2, 3):
        E = PyWaveEngine(1, nt, 0.001, 15.0, 20, 1000.0, 5000.0, nvx=100, nvz=100)
        assert E.streamed
        E.set_band_split(bands)
        E.set_forward_time_block(fwd_block)
        E.set_sources(np.array([[120.0, 10.0]], np.float32))
        rng = np.random.default_rng(10 + nt)
        obs = rng.standard_normal(nt * 100).astype(np.float32)
        vel = np.full((100, 100), 2200.0, np.float32)
        loss, grad = E.evaluate_adjoint(vel.reshape(-1), obs)
        rl, rg = orc.evaluate_adjoint_f32(orc.Geometry(nvx=100, nvz=100, nt=nt, dt=0.001), E.sources(), vel,
                                          E.wavelet(), obs.reshape(1, nt, 100))
        assert loss == rl
        assert np.array_equal(grad, rg)


def test_empty_batches_are_no_ops():
    from wgnet_amd import _lib
    from wgnet_amd.physics import WavePhysics, _dev_ptr, _stream

    ph = WavePhysics(shots=2, nt=50, dt=0.001)
    lib, h = ph.engine._lib, ph.engine.handle
    d = torch.zeros(16, device="cuda")
    ic = torch.zeros(16, dtype=torch.int32, device="cuda")
    u8 = torch.ones(16, dtype=torch.uint8, device="cuda")
    e64 = torch.zeros(4, dtype=torch.float64, device="cuda")
    P, S = _dev_ptr, _stream()
    calls = [
        lib.wg_simulate_dev(h, 0, P(d), P(d), S),
        lib.wg_loss_grad_dev(h, 0, P(d), P(d), P(d), P(d), None, P(e64), S),
        lib.wg_loss_grad_wavelets_dev(h, 0, P(d), P(d), P(d), P(d), P(d), None, P(e64), S),
        lib.wg_adjoint_apply_dev(h, 0, P(d), P(d), P(d), None, None, S),
        lib.wg_born_dev(h, 0, P(d), P(d), P(d), S),
        lib.wg_simulate_wavelets_dev(h, 0, P(d), P(d), P(d), S),
        lib.wg_loss_grad_units_dev(h, 0, P(ic), P(d), P(u8), P(d), P(d), P(d), P(d), P(d), None, P(e64), S),
        lib.wg_simulate_units_dev(h, 0, P(ic), P(d), P(u8), P(d), P(d), P(d), S),
        lib.wg_illumination_dev(h, 0, P(d), P(d), S),
    ]
    assert calls == [0] * len(calls)
    assert lib.wg_simulate_dev(h, -1, P(d), P(d), S) == _lib.WG_E_INVALID
