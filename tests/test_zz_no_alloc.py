"""wg_loss_noalloc_dev, the loss with no device allocation (its working data in
the library's fixed per-device scratch), against wg_loss_dev on the same
inputs through the C-ABI: two calls in a row, the second with a smaller batch
of other models, with and without a caller energy buffer and with per-model
wavelets; losses, gathers and energies bit for bit.  A batch beyond the
scratch is refused with the limit in the message."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SURFACE_SRC = np.array([[0, 10], [170, 10], [340, 10], [520, 10], [690, 10]], np.float32)


def test_noalloc_loss_equals_loss_dev():
    from wgnet_amd import _lib, synthetic
    from wgnet_amd.physics import WavePhysics, _stream

    S, nt = 5, 1000
    ph = WavePhysics(shots=S, nt=nt, dt=0.001)
    ph.engine.set_sources(SURFACE_SRC)
    lib, h = ph.engine._lib, ph.engine.handle
    R = ph.engine.nrec
    start, true = synthetic.batch(40, seed=777)
    obs = ph.simulate_torch(torch.from_numpy(true / 1000.0).cuda())
    vel = torch.from_numpy(start).cuda()  # m/s, as the C-ABI takes it
    P = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
    rng = np.random.default_rng(3)

    def call(fn, v, o, wav, with_energy):
        n = v.shape[0]
        loss = torch.full((n,), -1.0, device="cuda")
        tr = torch.full((n, S, nt, R), -1.0, device="cuda")
        en = torch.full((n,), -1.0, dtype=torch.float64, device="cuda") if with_energy else None
        _lib.check(fn(h, n, P(v), P(wav), P(o), P(loss), P(tr), P(en), _stream()))
        torch.cuda.synchronize()
        return loss, tr, en

    for sl in (slice(0, 40), slice(31, 40)):  # the second call is smaller, other models
        v, o = vel[sl].contiguous(), obs[sl].contiguous()
        n = v.shape[0]
        wav = torch.from_numpy(rng.standard_normal((n, S, nt)).astype(np.float32)).cuda()
        for w in (None, wav):
            l_ref, tr_ref, e_ref = call(lib.wg_loss_dev, v, o, w, True)
            for with_energy in (False, True):
                l_na, tr_na, e_na = call(lib.wg_loss_noalloc_dev, v, o, w, with_energy)
                print(f"{n} models, {'own' if w is not None else 'engine'} wavelets, caller energy {with_energy}: "
                      f"losses equal {bool(torch.equal(l_na, l_ref))}, gathers equal "
                      f"{bool(torch.equal(tr_na, tr_ref))}", flush=True)
                assert bool(torch.isfinite(l_ref).all()) and bool((l_ref > 0).all())
                assert torch.equal(l_na, l_ref)
                assert torch.equal(tr_na, tr_ref)
                if with_energy:
                    assert torch.equal(e_na, e_ref)

    # beyond the fixed scratch: refused before anything runs
    loss = torch.empty(40, device="cuda")
    tr = torch.empty((40, S, nt, R), device="cuda")
    rc = lib.wg_loss_noalloc_dev(h, 4000, P(vel), None, P(obs), P(loss), P(tr), None, _stream())
    msg = lib.wg_last_error().decode()
    print(f"4000 models: rc {rc}: {msg}")
    assert rc == _lib.WG_E_TOO_LARGE and "fixed scratch" in msg
