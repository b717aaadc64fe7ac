"""Pixel-grid FWI (fwi_eval.py:29-63) on the GPU at the reference's settings
(WavePhysics defaults: 5 shots, 2001 steps of 0.5 ms; bands 5, 8, 12, full;
gradient smoothing 1.5 cells; bounds [1.4, 4.6] km/s), with a small budget:
- `fwi` equals the reference's loop restated over the engine's one-model
  host calls (physics.py:66-76's evaluate_adjoint loop), bit for bit;
- `fwi_batch` gives each problem exactly what `fwi` gives it alone, with the
  optimiser side in worker processes and in threads of this process;
- the first objective value and gradient (start model, 5 Hz) equal the CPU
  oracle's on the same band-limited data;
- FWI lowers the full-band misfit."""

from __future__ import annotations

import numpy as np
import pytest
import torch
from scipy.ndimage import gaussian_filter
from scipy.optimize import minimize

pytestmark = pytest.mark.gpu

BANDS = [5.0, 8.0, 12.0, None]


def _reference_fwi(start_km, observed_full, physics, bands, budget, smooth_grad=1.5):
    """fwi_eval.py:32-57 with physics.loss_and_grad's per-model host loop."""
    N = 70
    x = start_km.astype(np.float64).reshape(-1)
    per_band = max(2, budget // len(bands))
    history = []
    for band in bands:
        physics.set_band(band)
        observed = physics.bandlimit(observed_full)
        calls = 0

        def objective(vec):
            nonlocal calls
            calls += 1
            vel = np.ascontiguousarray(vec.reshape(N, N).astype(np.float32).reshape(-1) * 1000.0, dtype=np.float32)
            loss, g = physics.engine.evaluate_adjoint(vel, np.ascontiguousarray(observed.reshape(-1), np.float32))
            g = (g * 1000.0).astype(np.float32)
            g = gaussian_filter(g, smooth_grad, mode="nearest") if smooth_grad > 0 else g
            return float(np.float32(loss)), g.reshape(-1).astype(np.float64)

        res = minimize(objective, x, jac=True, method="L-BFGS-B", bounds=[(1.4, 4.6)] * (N * N),
                       options={"maxfun": per_band, "maxiter": per_band, "maxcor": 10})
        x = res.x
        history.append({"band_hz": band, "misfit": float(res.fun), "solves": calls})
    physics.set_band(None)
    return x.reshape(N, N).astype(np.float32), history


def _problems(n, seed):
    from wgnet_amd import synthetic
    from wgnet_amd.fwi_eval import macro_start

    start, true = synthetic.batch(n, seed=seed)
    truth = (true / 1000.0).astype(np.float32)
    linear1d = np.repeat((start / 1000.0).mean(axis=(0, 2))[None, :, None], 70, axis=2).astype(np.float32)
    starts = [linear1d[0], macro_start(truth[:1], 6.0)[0], linear1d[0]]
    return truth, np.stack(starts)


def test_fwi_equals_reference_loop_and_batch_equals_fwi():
    from oracle import oracle as O
    from wgnet_amd.fwi_eval import fwi, fwi_batch, misfit_full
    from wgnet_amd.physics import WavePhysics

    phys = WavePhysics()
    truth, starts = _problems(2, seed=41)
    tmodels = np.stack([truth[0], truth[0], truth[1]])
    observed = phys.simulate(tmodels)
    budget = 12  # 3 solves per band

    # the first objective value against the CPU oracle
    phys.set_band(5.0)
    obs5 = phys.bandlimit(observed[0])
    l5, g5 = phys.loss_and_grad(starts[:1], obs5[None])
    g = O.Geometry(nt=2001, dt=0.0005)
    rl, rg = O.evaluate_adjoint_f32(g, phys.engine.sources(), np.ascontiguousarray(starts[0] * 1000.0, np.float32),
                                    phys.engine.wavelet(), obs5)
    assert rl == float(l5[0]) and (rg * np.float32(1000.0)).tobytes() == g5[0].tobytes()
    phys.set_band(None)

    f_ref, h_ref = _reference_fwi(starts[0], observed[0], phys, BANDS, budget)
    f_one, h_one = fwi(starts[0], observed[0], phys, BANDS, budget)
    print("fwi history", h_one, flush=True)
    assert f_one.tobytes() == f_ref.tobytes() and h_one == h_ref
    assert phys.fmax is None

    finals, hists, calls = fwi_batch(starts, observed, phys, BANDS, budget)
    print("batched calls", calls, [sum(h["solves"] for h in hh) for hh in hists], flush=True)
    assert finals[0].tobytes() == f_one.tobytes() and hists[0] == h_one
    for p in (1, 2):
        f_p, h_p = fwi(starts[p], observed[p], phys, BANDS, budget)
        assert finals[p].tobytes() == f_p.tobytes() and hists[p] == h_p
    f0, h0, _ = fwi_batch(starts, observed, phys, BANDS, budget, workers=0)  # threads of this process
    assert all(f0[p].tobytes() == finals[p].tobytes() and h0[p] == hists[p] for p in range(3))
    for p in range(3):
        assert misfit_full(finals[p], observed[p], phys) < misfit_full(starts[p], observed[p], phys)
        assert finals[p].min() >= 1.4 and finals[p].max() <= 4.6
