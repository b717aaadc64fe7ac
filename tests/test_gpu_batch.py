"""Uneven batches (grids, step counts, shot counts, source positions and
passbands mixed in one HIP graph) against standalone engines, one per solve.

Padding slots contribute exact zeros to every sum (thread-strided energy,
256-chunk residual tree, shot sum of the fold), so each solve's loss and
gradient must equal its standalone engine's bit for bit.  The batch runs as
one engine call with one forward and one adjoint launch for all groups (the
default), or group by group on several streams; both give the same bits, also
when a grid without compiled kernels makes the one call run its groups one
after another."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _solves():
    from wgnet_amd.batch import Solve

    return [
        Solve(70, 70, 300, np.array([0.0, 345.5, 690.0], np.float32), 9.0),
        Solve(70, 70, 300, np.array([120.0], np.float32), None),
        Solve(70, 70, 400, np.array([10.0, 200.0, 333.3, 500.0, 680.0], np.float32), 12.0),
        Solve(100, 100, 300, np.array([50.0, 990.0], np.float32), 6.0),
        Solve(100, 100, 300, np.array([400.0, 410.0, 777.0, 901.0], np.float32), 15.0),
        Solve(70, 70, 300, np.array([690.0, 1.5], np.float32), 15.0),
        Solve(141, 141, 300, np.array([100.0, 700.0, 1395.0], np.float32), 12.0),
    ]


def _models(solves, seed):
    from wgnet_amd import synthetic

    start, true = [], []
    for i, s in enumerate(solves):
        a, b = synthetic.batch(1, seed=seed + i, nvz=s.nvz, nvx=s.nvx)
        start.append(a[0])
        true.append(b[0])
    return start, true


def test_uneven_batch_equals_standalone_engines():
    from wgnet_amd.batch import UnevenBatch, _bandlimit
    from wgnet_amd.physics import _dev_ptr, _stream, lowpass_kernel
    from wgnet_amd import _lib
    from wgnet_amd.wave_gauss_rs import PyWaveEngine

    solves = _solves()
    ub = UnevenBatch(solves, dt=0.001)
    tiles = ub.autotune(repeats=1)
    print(f"autotuned tiles: {tiles}")
    start, true = _models(solves, 500)
    ub.set_models(true)
    obs = ub.simulate(full_band=True)
    ub.set_observed(obs)
    ub.set_models(start)
    ub.run_eager()
    torch.cuda.synchronize()
    eager = [(l.clone(), g.clone()) for l, g in ub.results()]
    for i, s in enumerate(solves):
        e = PyWaveEngine(s.shots, s.nt, 0.001, 15.0, 20, 1000.0, 5000.0, nvx=s.nvx, nvz=s.nvz)
        e.set_sources(np.stack([s.src_x, np.full(s.shots, 10.0, np.float32)], 1).astype(np.float32))
        w = e.wavelet()  # full-band Ricker
        # observed: the standalone engine's own full-band gathers of the true model, low-passed
        full = torch.from_numpy(e.simulate(np.ascontiguousarray(true[i].reshape(-1)))).cuda()
        if s.band_hz is not None:
            w = np.convolve(w, lowpass_kernel(s.band_hz, 0.001), mode="full")[: s.nt].astype(np.float32)
        assert torch.equal(full, obs[i]), f"solve {i}: full-band gathers differ"
        d = full if s.band_hz is None else _bandlimit(full, s.band_hz, 0.001)
        e.set_wavelet(np.ascontiguousarray(w))
        vel = torch.from_numpy(np.ascontiguousarray(start[i][None])).cuda()
        loss = torch.empty(1, device="cuda")
        grad = torch.empty((1, s.nvz, s.nvx), device="cuda")
        en = torch.empty(1, dtype=torch.float64, device="cuda")
        _lib.check(e._lib.wg_loss_grad_dev(e.handle, 1, _dev_ptr(vel), _dev_ptr(d.contiguous()), _dev_ptr(loss),
                                           _dev_ptr(grad), None, _dev_ptr(en), _stream()))
        torch.cuda.synchronize()
        print(f"solve {i} ({s.nvx}x{s.nvz}, nt {s.nt}, {s.shots} shots, band {s.band_hz}): "
              f"loss {float(eager[i][0]):.6e} vs {float(loss[0]):.6e}")
        assert torch.equal(eager[i][0], loss[0])
        assert torch.equal(eager[i][1], grad[0])


def test_uneven_batch_graph_replay_equals_eager():
    """The captured graph, replayed three times with eager work and other
    models in between, reproduces the eager results each time."""
    from wgnet_amd.batch import UnevenBatch

    solves = _solves()
    ub = UnevenBatch(solves, dt=0.001)
    start, true = _models(solves, 600)
    ub.set_models(true)
    ub.set_observed(ub.simulate(full_band=True))
    ub.set_models(start)
    ub.run_eager()
    torch.cuda.synchronize()
    ref = [(l.clone(), g.clone()) for l, g in ub.results()]
    ub.capture()
    other, _ = _models(solves, 700)
    for rep in range(3):
        ub.set_models(other)
        ub.run_eager()  # eager work between replays
        torch.cuda.synchronize()
        ub.set_models(start)
        ub.replay()
        torch.cuda.synchronize()
        for i, (l, g) in enumerate(ub.results()):
            assert torch.equal(l, ref[i][0]), f"replay {rep} solve {i}: loss"
            assert torch.equal(g, ref[i][1]), f"replay {rep} solve {i}: grad"


@pytest.mark.parametrize("extra_grid", [None, 90])
def test_fused_call_equals_per_group_streams(extra_grid):
    from wgnet_amd.batch import Solve, UnevenBatch

    solves = _solves()
    if extra_grid:  # no compiled kernels for this grid: the fused call runs the groups one by one
        solves.append(Solve(extra_grid, extra_grid, 250, np.array([33.0, 444.0], np.float32), 9.0))
    start, true = _models(solves, 800)
    res = {}
    for fused in (True, False):
        ub = UnevenBatch(solves, dt=0.001, fused=fused)
        ub.set_models(true)
        ub.set_observed(ub.simulate(full_band=True))
        ub.set_models(start)
        ub.run_eager()
        ub.capture()
        ub.replay()
        torch.cuda.synchronize()
        res[fused] = [(l.clone(), g.clone()) for l, g in ub.results()]
        del ub
    for i, ((l1, g1), (l0, g0)) in enumerate(zip(res[True], res[False])):
        assert bool(torch.isfinite(l1)) and float(l1) > 0.0, i
        assert torch.equal(l1, l0) and torch.equal(g1, g0), (extra_grid, i)
