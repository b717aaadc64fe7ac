"""The uneven batch at its stated size: 512 solves, each drawing its grid from
{70 x 70, 100 x 100, 141 x 141}, its step count from {1000, 1500, 2000}, 1-5
shots at random surface positions and a passband from {6, 9, 12, 15} Hz
applied to both wavelet and data; the same draw as bench.py's uneven workload.

Checked:
  - one solve of every (grid, nt) group (the one with the most shots) against
    the float32 oracle run on the same inputs (its band-limited wavelet, its
    band-limited observed gathers): loss |rel| <= 1e-6, gradient rel-L2 <=
    1e-6; and that solve's full-band observed gathers against the oracle's
    forward of the true model, rel-L2 <= 1e-5;
  - all 512 solves: finite losses >= 0 and finite gradients, at least one
    positive loss per group; the captured HIP graph replayed twice, with eager
    work between, equal to the eager run bit for bit.
The batch runs as one launch (UnevenBatch's default): every unit or row band
a task of one kernel, forward, residual and adjoint."""

from __future__ import annotations

import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_uneven_512_solves_against_oracle(orc):
    from wgnet_amd import synthetic
    from wgnet_amd.batch import UnevenBatch, random_solves
    from wgnet_amd.physics import lowpass_kernel

    dt = 0.001
    solves = random_solves(512, seed=2000)
    assert {s.nvx for s in solves} == {70, 100, 141} and {s.nt for s in solves} == {1000, 1500, 2000}
    ub = UnevenBatch(solves, dt=dt)
    start, true = [], []
    for i, s in enumerate(solves):
        a, b = synthetic.batch(1, seed=2000 + 7919 * i, nvz=s.nvz, nvx=s.nvx)
        start.append(a[0])
        true.append(b[0])
    ub.set_models(true)
    full = ub.simulate(full_band=True)
    ub.set_observed(full)
    ub.set_models(start)
    ub.run_eager()
    torch.cuda.synchronize()
    eager = [(l.clone(), g.clone()) for l, g in ub.results()]
    losses = torch.stack([l for l, _ in eager])
    assert bool(torch.isfinite(losses).all()) and bool((losses >= 0).all())
    assert all(bool(torch.isfinite(g).all()) for _, g in eager)

    ub.capture()
    for rep in range(2):
        ub.set_models(true)
        ub.run_eager()  # other work between replays
        ub.set_models(start)
        ub.replay()
        torch.cuda.synchronize()
        for i, (l, g) in enumerate(ub.results()):
            assert torch.equal(l, eager[i][0]) and torch.equal(g, eager[i][1]), (rep, i)

    # one solve per (grid, nt) group against the oracle
    picks = []
    for g in ub.groups:
        i = max(g.ids, key=lambda k: solves[k].shots)
        b = g.ids.index(i)
        assert float(torch.stack([eager[k][0] for k in g.ids]).max()) > 0.0
        picks.append((i, g.buf["obs"][b, : solves[i].shots].cpu().numpy()))

    def one(item):
        i, obs_band = item
        s = solves[i]
        w = orc.ricker_f32(s.nt, dt, 15.0)
        wb = np.convolve(w, lowpass_kernel(s.band_hz, dt), mode="full")[: s.nt].astype(np.float32)
        src = np.stack([s.src_x, np.full(s.shots, s.src_z, np.float32)], 1).astype(np.float32)
        geo = orc.Geometry(nvx=s.nvx, nvz=s.nvz, nt=s.nt, dt=dt)
        rl, rg = orc.evaluate_adjoint_f32(geo, src, start[i], wb, obs_band)
        x0 = float(s.src_x[0])
        tr = orc.forward_f32(orc.Geometry(nvx=s.nvx, nvz=s.nvz, nt=s.nt, dt=dt, src_x=x0, src_z=s.src_z), true[i], w)
        return i, rl, rg, tr

    threads = max(1, min(16, len(os.sched_getaffinity(0)), len(picks)))
    with ThreadPoolExecutor(threads) as pool:
        res = list(pool.map(one, picks))
    for i, rl, rg, tr in res:
        s = solves[i]
        loss, grad = float(eager[i][0]), eager[i][1].cpu().numpy()
        e_g = relerr(grad, rg)
        e_tr = relerr(full[i][0].cpu().numpy(), tr)
        print(f"solve {i}: {s.nvx}x{s.nvz}, nt {s.nt}, {s.shots} shots, {s.band_hz} Hz: loss {loss:.9e} vs "
              f"{rl:.9e}, gradient rel-L2 {e_g:.2e}, full-band gathers rel-L2 {e_tr:.2e}", flush=True)
        assert abs(loss - rl) <= 1e-6 * max(rl, 1e-30)
        assert e_g <= 1e-6
        assert e_tr <= 1e-5
