"""The wave-cache writer's format and bookkeeping (data.py:137-179) on the
CPU, with a deterministic stand-in for the solver: the file is the
reference loop's byte for byte (one model at a time, [:, ::8, :] decimation,
float16 cast, float64 square sums in model order), progress is written at
every 500th model and resumed from, and done.json holds the RMS."""

from __future__ import annotations

import json

import numpy as np
import torch

from wgnet_amd import wavecache


class _FakePhysics:
    """simulate_torch of a batch = the per-model function of each field."""

    def __init__(self, shots=5, nt=41):
        self.shots, self.nt = shots, nt

    def simulate_torch(self, fields_km):
        f = fields_km.float()
        B = f.shape[0]
        t = torch.arange(self.nt, dtype=torch.float32)[None, None, :, None]
        s = torch.arange(self.shots, dtype=torch.float32)[None, :, None, None]
        r = torch.arange(70, dtype=torch.float32)[None, None, None, :]
        m = f.mean((1, 2)).view(B, 1, 1, 1)
        return torch.sin(0.37 * t * m + 0.11 * s + 0.05 * r) * (1.0 + 0.1 * f[:, :1, :70].view(B, 1, 1, 70))


def _reference_loop(phys, fields):
    """data.py:159-171 restated with the stand-in solver."""
    nt_dec = (phys.nt - 1) // 8 + 1
    out = np.zeros((len(fields), phys.shots, nt_dec, 70), np.float16)
    sq_sum = 0.0
    for i in range(len(fields)):
        traces = phys.simulate_torch(torch.from_numpy(fields[i:i + 1]))[0].numpy()[:, ::8, :]
        sq_sum += float(np.square(traces, dtype=np.float64).sum())
        out[i] = traces.astype(np.float16)
    rms = float(np.sqrt(sq_sum / (len(fields) * phys.shots * nt_dec * 70)))
    return out, rms


def _fields(n, seed):
    rng = np.random.default_rng(seed)
    return (1.5 + 3.0 * rng.random((n, 70, 70))).astype(np.float32)


def test_cache_equals_reference_loop_and_reopens(tmp_path):
    phys = _FakePhysics()
    fields = {"train": _fields(11, 0), "val": _fields(3, 1)}
    mms = wavecache.ensure_wave_cache(fields, tmp_path, ("train", "val"), phys, batch=4, log=lambda *_: None)
    for split in ("train", "val"):
        ref, rms = _reference_loop(phys, fields[split])
        assert mms[split].shape == ref.shape and mms[split].dtype == np.float16
        assert mms[split].tobytes() == ref.tobytes()
        done = json.loads((tmp_path / f"waves_{split}.done.json").read_text())
        assert done == {"split": split, "count": len(fields[split]), "rms": rms}
        assert not (tmp_path / f"waves_{split}.progress.json").exists()
    assert wavecache.rms_scale(tmp_path) == max(json.loads((tmp_path / "waves_train.done.json").read_text())["rms"], 1e-6)
    # a second call opens the finished files without simulating
    class _Fail:
        shots, nt = phys.shots, phys.nt

        def simulate_torch(self, f):
            raise AssertionError("simulated again")
    again = wavecache.ensure_wave_cache(fields, tmp_path, ("train", "val"), _Fail(), log=lambda *_: None)
    assert again["train"].tobytes() == mms["train"].tobytes()


def test_progress_every_n_models_and_resume(tmp_path):
    """An interrupted run leaves progress at the last multiple of the
    progress interval; the next run resumes there and the file matches an
    uninterrupted run."""
    phys = _FakePhysics()
    fields = {"train": _fields(9, 2)}

    class _Stop(Exception):
        pass

    class _Interrupting(_FakePhysics):
        calls = 0

        def simulate_torch(self, f):
            _Interrupting.calls += 1
            if _Interrupting.calls == 4:
                raise _Stop()
            return super().simulate_torch(f)

    try:
        wavecache.ensure_wave_cache(fields, tmp_path, ("train",), _Interrupting(), batch=2, progress_every=3,
                                    log=lambda *_: None)
    except _Stop:
        pass
    prog = json.loads((tmp_path / "waves_train.progress.json").read_text())
    # batches of 2 models, each simulated while the one before it is written:
    # the 4th batch's simulation fails after models 0..3 are written (progress at 3)
    assert prog["completed"] == 3
    sq3 = 0.0
    for i in range(3):  # the float64 square sums of models 0..2, in order
        tr = phys.simulate_torch(torch.from_numpy(fields["train"][i:i + 1]))[0].numpy()[:, ::8, :]
        sq3 += float(np.square(tr, dtype=np.float64).sum())
    assert prog["sq_sum"] == sq3
    mm = wavecache.ensure_wave_cache(fields, tmp_path, ("train",), phys, batch=2, progress_every=3,
                                     log=lambda *_: None)["train"]
    full, rms_full = _reference_loop(phys, fields["train"])
    assert mm.tobytes() == full.tobytes()
    assert json.loads((tmp_path / "waves_train.done.json").read_text())["rms"] == rms_full


def test_rms_scale_simulates_missing_train_split(tmp_path):
    """data.py:183-185: rms_scale builds the train cache when it is missing."""
    phys = _FakePhysics()
    fields = {"train": _fields(5, 3)}
    try:
        wavecache.rms_scale(tmp_path)
        raise AssertionError("expected FileNotFoundError")
    except FileNotFoundError:
        pass
    rms = wavecache.rms_scale(tmp_path, fields, phys)
    assert rms == max(_reference_loop(phys, fields["train"])[1], 1e-6)
    assert wavecache.wave_path(tmp_path, "train").exists()


def test_rejects_empty_batches(tmp_path):
    import pytest

    with pytest.raises(ValueError):
        wavecache.ensure_wave_cache({"train": _fields(2, 4)}, tmp_path, ("train",), _FakePhysics(), batch=0)
