"""The wave-cache writer on the GPU against the reference loop
(data.py:159-171: engine.simulate of one model at a time, [:, ::8, :],
float16, float64 square sums in model order) run through the same engine's
one-model host path: byte-identical file and identical RMS, at the
reference's simulation settings (5 shots, 15 Hz, dt 0.5 ms, 2001 steps);
the first model is also checked against the CPU oracle's forward."""

from __future__ import annotations

import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_wave_cache_equals_per_model_reference_loop(tmp_path):
    from wgnet_amd import synthetic, wavecache

    start, true = synthetic.batch(7, seed=515)
    fields = {"train": (true / 1000.0).astype(np.float32)}
    phys = wavecache.make_physics()
    mm = wavecache.ensure_wave_cache(fields, tmp_path, ("train",), phys, batch=3, log=lambda *_: None)["train"]
    assert mm.shape == (7, 5, 251, 70)
    ref = np.zeros(mm.shape, np.float16)
    sq_sum = 0.0
    for i in range(7):
        vel = np.ascontiguousarray(fields["train"][i].reshape(-1) * 1000.0, dtype=np.float32)
        traces = phys.engine.simulate(vel)[:, ::8, :]
        sq_sum += float(np.square(traces, dtype=np.float64).sum())
        ref[i] = traces.astype(np.float16)
    assert mm.tobytes() == ref.tobytes()
    rms = float(np.sqrt(sq_sum / (7 * 5 * 251 * 70)))
    assert json.loads((tmp_path / "waves_train.done.json").read_text())["rms"] == rms
    assert np.abs(mm.astype(np.float32)).max() > 0.0
    # and the first model's gathers are the CPU oracle's, decimated and cast
    from oracle import oracle as O

    g = O.Geometry(nt=2001, dt=0.0005)
    w = O.ricker_f32(2001, 0.0005, 15.0)
    src = phys.engine.sources()
    vel = np.ascontiguousarray(np.repeat(fields["train"][:1].reshape(1, -1) * 1000.0, 5, 0), dtype=np.float32)
    cpu = O.batch_forward_f32(g, src, vel, w, 5)[:, ::8, :]
    assert cpu.astype(np.float16).tobytes() == mm[0].tobytes()
