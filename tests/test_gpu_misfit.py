"""The misfit as a PyTorch loss: the adjoint runs only when a gradient is
needed, the zero-energy error of the reference (wave_wgpu.rs:853 asserts
energy > 0, which NaN fails too), and a work-queue wait that gives up."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _timing(lib, h):
    fms, fn, ams, an = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double(), ctypes.c_int64()
    lib.wg_timing_read(h, ctypes.byref(fms), ctypes.byref(fn), ctypes.byref(ams), ctypes.byref(an))
    return fn.value, an.value


@pytest.fixture(scope="module")
def setup():
    from wgnet_amd import synthetic
    from wgnet_amd.physics import WavePhysics

    ph = WavePhysics(shots=3, nt=300, dt=0.001)
    start, true = synthetic.batch(4, seed=515)
    obs = ph.simulate_torch(torch.from_numpy(true / 1000.0).cuda())
    f = torch.from_numpy(start / 1000.0).cuda()
    return ph, f, obs


def test_misfit_without_autograd_runs_no_adjoint(setup):
    """misfit() under no_grad, or on fields that need no gradient, launches
    the forward only; its losses equal the loss-and-gradient call's."""
    ph, f, obs = setup
    lib, h = ph.engine._lib, ph.engine.handle
    ref, _ = ph.loss_and_grad_torch(f, obs)
    fr = f.clone().requires_grad_(True)
    lib.wg_timing_enable(h, 1)
    _timing(lib, h)
    with torch.no_grad():
        l1 = ph.misfit(fr, obs)
    torch.cuda.synchronize()
    n1 = _timing(lib, h)
    l2 = ph.misfit(f, obs)  # no requires_grad
    torch.cuda.synchronize()
    n2 = _timing(lib, h)
    l3 = ph.misfit(fr, obs)
    l3.sum().backward()
    torch.cuda.synchronize()
    n3 = _timing(lib, h)
    lib.wg_timing_enable(h, 0)
    print(f"launches (forward, adjoint): no_grad {n1}, detached {n2}, with autograd {n3}")
    assert n1[0] >= 1 and n1[1] == 0
    assert n2[0] >= 1 and n2[1] == 0
    assert n3[1] >= 1
    assert not l1.requires_grad and not l2.requires_grad
    assert torch.equal(l1, ref) and torch.equal(l2, ref) and torch.equal(l3.detach(), ref)
    _, g_ref = ph.loss_and_grad_torch(f, obs)
    assert torch.equal(fr.grad, g_ref)


def test_zero_or_nan_energy_raises(setup):
    ph, f, obs = setup
    bad = obs.clone()
    bad[1] = 0.0
    with pytest.raises(ValueError, match="zero energy"):
        ph.loss_and_grad_torch(f, bad)
    with pytest.raises(ValueError, match="zero energy"):
        ph.loss_torch(f, bad)
    nan = obs.clone()
    nan[2, 0, 5, 7] = float("nan")
    with pytest.raises(ValueError, match="zero energy"):
        ph.loss_and_grad_torch(f, nan)
    # a tensor that passed once is checked again after an in-place change
    ok = obs.clone()
    ph.loss_and_grad_torch(f, ok)
    ph.loss_and_grad_torch(f, ok)
    ok[0].zero_()
    with pytest.raises(ValueError, match="zero energy"):
        ph.loss_and_grad_torch(f, ok)


def test_queue_wait_that_gives_up_poisons_losses_and_gradients():
    """wg_set_queue_wait_limit(-1) makes every work-queue wait give up at once
    (every kernel still ends normally): each loss and every gradient value of
    that call is NaN.  The default limit gives the finite results back."""
    from wgnet_amd import synthetic
    from wgnet_amd.physics import WavePhysics

    B, S, nt = 64, 5, 300
    ph = WavePhysics(shots=S, nt=nt, dt=0.001)
    lib, h = ph.engine._lib, ph.engine.handle
    start, true = synthetic.batch(B, seed=616)
    obs = ph.simulate_torch(torch.from_numpy(true / 1000.0).cuda())
    f = torch.from_numpy(start / 1000.0).cuda()
    l0, g0 = ph.loss_and_grad_torch(f, obs)
    assert lib.wg_set_queue_wait_limit(h, -1) == 0
    l1, g1 = ph.loss_and_grad_torch(f, obs)
    torch.cuda.synchronize()
    assert bool(torch.isnan(l1).all()) and bool(torch.isnan(g1).all())
    assert lib.wg_set_queue_wait_limit(h, 1 << 24) == 0
    l2, g2 = ph.loss_and_grad_torch(f, obs)
    assert torch.equal(l2, l0) and torch.equal(g2, g0)
