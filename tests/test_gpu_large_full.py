"""The HBM-streamed large grid at its stated size: one model of 1000 x 300
cells (1044 x 344 padded), 20 shots evenly spaced on the surface, 1000
receivers, 4000 steps of 0.8 ms, with the adjoint checkpointed every 100 steps
and recomputing each segment (the snapshot budget is set below one model's
snapshots, 113 GB, so the engine must checkpoint).

Checked:
  - the FP64-accumulated dot-product test of the wavelet operator F (source
    wavelet -> gathers) through the checkpointed adjoint, y = F x, for x a
    random wavelet low-passed to the 15 Hz Ricker's band and for white-noise
    x: |<F x, y> - <x, F^T y>| <= 1e-5 |<F x, y>|.  The gate is float32
    roundoff over 4000 steps, not transposition: the float32 oracle pair (the
    reference's float32 forward restated, and the WGSL adjoint) shows the
    same growth with nt on a small grid, 4e-7 at nt = 500 to 9e-6 at nt =
    4000, where its float64 version stays at 1e-15
    (tests/test_oracle.py::test_dot_product_float32_level_grows_with_nt); the
    GPU measures 1.9e-6 here with the band-limited x, and 2.6e-7 at
    70 x 70 x 1000;
  - the loss and gradient of the checkpointed path: finite, gradient nonzero;
  - checkpointed = every snapshot stored, bit for bit, on the same grid and
    shots with nt shortened to 400 (four segments)."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

NVX, NVZ, S, DT = 1000, 300, 20, 0.0008


def _physics(nt):
    from wgnet_amd.physics import WavePhysics

    ph = WavePhysics(shots=S, nt=nt, dt=DT, nvx=NVX, nvz=NVZ)
    assert ph.engine.streamed and ph.engine._lib.wg_engine_bands(ph.engine.handle) == 0
    src = np.stack([np.linspace(0.0, (NVX - 1) * 10.0, S), np.full(S, 10.0)], 1).astype(np.float32)
    ph.engine.set_sources(src)
    return ph


def test_large_dot_product_checkpointed():
    from wgnet_amd import _lib, synthetic
    from wgnet_amd.physics import _stream, lowpass_kernel

    nt = 4000
    ph = _physics(nt)
    lib, h = ph.engine._lib, ph.engine.handle
    assert lib.wg_set_snapshot_budget(h, ctypes.c_uint64(16 << 30)) == 0  # below one model's 113 GB
    start, _ = synthetic.large_batch(1, seed=3000, nvz=NVZ, nvx=NVX)
    vel = torch.from_numpy(start).cuda()
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    rng = np.random.default_rng(31)
    taps = lowpass_kernel(15.0, DT)
    for label, smooth in (("15 Hz band-limited x", True), ("white-noise x", False)):
        x = rng.standard_normal((1, S, nt)).astype(np.float32)
        if smooth:
            x = np.stack([[np.convolve(x[0, k], taps, mode="same") for k in range(S)]]).astype(np.float32)
        xd = torch.from_numpy(x).cuda()
        fx = torch.empty((1, S, nt, NVX), device="cuda")
        _lib.check(lib.wg_simulate_wavelets_dev(h, 1, P(vel), P(xd), P(fx), _stream()))
        ftfx = torch.empty((1, S, nt), device="cuda")
        _lib.check(lib.wg_adjoint_apply_dev(h, 1, P(vel), P(fx), None, P(ftfx), None, _stream()))
        lhs, rhs = float((fx.double() ** 2).sum()), float((xd.double() * ftfx.double()).sum())
        rel = abs(lhs - rhs) / abs(lhs)
        print(f"1000x300x4000, 20 shots, checkpointed adjoint, {label}: <Fx,Fx>={lhs:.12e} <x,F^T F x>={rhs:.12e} "
              f"rel {rel:.3e}", flush=True)
        assert rel <= 1e-5
        del fx


def test_large_loss_and_gradient_checkpointed():
    from wgnet_amd import synthetic

    nt = 4000
    ph = _physics(nt)
    assert ph.engine._lib.wg_set_snapshot_budget(ph.engine.handle, ctypes.c_uint64(16 << 30)) == 0
    start, true = synthetic.large_batch(1, seed=3001, nvz=NVZ, nvx=NVX)
    obs = ph.simulate_torch(torch.from_numpy(true / 1000.0).cuda())
    l, g = ph.loss_and_grad_torch(torch.from_numpy(start / 1000.0).cuda(), obs)
    torch.cuda.synchronize()
    print(f"1000x300x4000, 20 shots: loss {float(l[0]):.6e}, |grad| {float(g.norm()):.6e}", flush=True)
    assert bool(torch.isfinite(l).all()) and float(l[0]) > 0.0
    assert bool(torch.isfinite(g).all()) and float(g.abs().max()) > 0.0


def test_large_checkpointed_equals_stored():
    from wgnet_amd import synthetic

    nt = 400
    ph = _physics(nt)
    lib, h = ph.engine._lib, ph.engine.handle
    start, true = synthetic.large_batch(1, seed=3002, nvz=NVZ, nvx=NVX)
    obs = ph.simulate_torch(torch.from_numpy(true / 1000.0).cuda())
    f = torch.from_numpy(start / 1000.0).cuda()
    assert lib.wg_set_snapshot_budget(h, ctypes.c_uint64(64 << 30)) == 0  # 11.3 GB of snapshots: stored
    l_st, g_st = ph.loss_and_grad_torch(f, obs)
    assert lib.wg_set_snapshot_budget(h, ctypes.c_uint64(4 << 30)) == 0  # checkpoints every 100 steps
    l_ck, g_ck = ph.loss_and_grad_torch(f, obs)
    torch.cuda.synchronize()
    print(f"nt 400: stored loss {float(l_st[0]):.9e}, checkpointed {float(l_ck[0]):.9e}", flush=True)
    assert torch.equal(l_st, l_ck) and torch.equal(g_st, g_ck)
    assert float(g_st.abs().max()) > 0.0
