"""The trained preconditioner inside refine: a small network trained on pairs
from the batched solver (at the scenes' initial fields), then plain and
preconditioned projected L-BFGS on held-out models at equal budgets.  The
test checks the invariants (budgets, descent within a band, fallback
accounting) and that the learned update steps lower the median misfit below
plain L-BFGS's at the same budget; the full-size measurement is
bench.py --workload refine."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_preconditioned_refine_against_plain():
    import bench
    from wgnet_amd.physics import WavePhysics
    from wgnet_amd.precond import Preconditioner, make_pairs, train
    from wgnet_amd.refine import refine_batch

    bands = (6.0, 9.0, 12.0, 15.0)
    ph = WavePhysics(shots=5, nt=1000, dt=0.001)
    ph.engine.set_sources(bench.SURFACE_SRC)
    offset = bench.refine_lattice_field()
    grads, fields, target = make_pairs(ph, 2048, seed=21, bands_hz=bands, offset_km=offset)
    net = Preconditioner(max_batch=256, seed=2)
    hist = train(net, grads, fields, target, epochs=30, lr=3e-3)
    zero = float((target ** 2).mean())
    print(f"training loss {hist[0]:.4f} -> {hist[-1]:.4f} (zero prediction {zero:.4f})")
    assert hist[-1] < zero  # the network explains part of the update

    start, true, gauss = bench.refine_scenes(16, seed=700000)
    obs = ph.simulate_torch(torch.from_numpy(true).cuda())
    budget = 40
    plain = refine_batch(start, gauss, obs, ph, budget_solves=budget, bands_hz=bands)
    pc = refine_batch(start, gauss, obs, ph, budget_solves=budget, bands_hz=bands, update=net.update_hook())
    pd = refine_batch(start, gauss, obs, ph, budget_solves=budget, bands_hz=bands, precondition=net.hook())
    mp = np.array([r.final_loss for r in plain])
    mq = np.array([r.final_loss for r in pc])
    init = np.array([r.initial_loss for r in plain])
    print(f"median misfit: initial {np.median(init):.4e}, plain {np.median(mp):.4e}, preconditioned "
          f"{np.median(mq):.4e}; preconditioned lower in {np.mean(mq < mp):.2f} of models")
    for r in plain + pc + pd:
        assert r.solves <= budget + 2
        # accepted steps never raise the misfit within a band
        for a, b in zip(r.history, r.history[1:]):
            if a["band_hz"] == b["band_hz"] and b["accepted"]:
                assert b["loss"] <= a["loss"]
    assert np.median(mq) < np.median(init)
    assert np.median(mq) < np.median(mp)  # the learned first steps pay for themselves
    assert any(h["jump"] for r in pc for h in r.history)
    assert any(h["preconditioned"] for r in pd for h in r.history)
