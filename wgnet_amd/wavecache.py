"""The simulated shot-gather cache that WGNet trains from, written at GPU speed.

Counterpart of OpenFWIData.ensure_wave_cache / rms_scale / wave_path
(python/wgnet/data.py:134-188).  The on-disk format and its bookkeeping are
the reference's:
  waves_<split>.f16.bin        float16 memmap [models, 5, 251, 70]: each
                               model's gathers (5 shots, 15 Hz Ricker, 0.5 ms
                               for 2001 steps) decimated to 4 ms ([:, ::8, :])
  waves_<split>.progress.json  {"completed", "sq_sum"} every 500 models, so
                               an interrupted run resumes where it stopped
  waves_<split>.done.json      {"split", "count", "rms"} when a split is done;
                               rms = sqrt(sum of the float64 squares of the
                               decimated float32 gathers / values)
What changes is the simulation: the reference calls engine.simulate once per
model (data.py:160-163); here a batch of models is one call of the batched
engine, decimated and cast to float16 on the device, and only the decimated
gathers cross PCIe; the next batch simulates while the host writes the
current one.  Per model the host forms the same float64 square sum (numpy,
in model order), and the device's float16 cast rounds to nearest even as
numpy's does, so the files are byte-identical.
"""

from __future__ import annotations

import json
import time
from pathlib import Path

import numpy as np
import torch

from .physics import DECIMATE, N, NT, NT_DEC, SHOTS, WavePhysics

SPLITS = ("train", "val", "test")
PROGRESS_EVERY = 500  # data.py:165


def wave_path(cache_dir: Path | str, split: str) -> Path:
    """data.py:134-135."""
    return Path(cache_dir) / f"waves_{split}.f16.bin"


def make_physics(shots: int = SHOTS, nt: int = NT) -> WavePhysics:
    """The engine of data.py:32-33 (5 shots at the default source layout,
    15 Hz Ricker, dt 0.5 ms, 2001 steps) behind the batched interface."""
    return WavePhysics(shots=shots, nt=nt)


def ensure_wave_cache(fields_km: dict, cache_dir: Path | str, splits: tuple[str, ...] = SPLITS, physics=None,
                      batch: int = 256, progress_every: int = PROGRESS_EVERY, log=print) -> dict:
    """Simulates (resumably) and opens the decimated float16 gather cache of
    each split (data.py:137-179).  fields_km maps split -> [models, 70, 70]
    km/s; returns split -> read-only memmap [models, 5, 251, 70].  `physics`
    is a WavePhysics (or anything with simulate_torch and shots/nt); `batch`
    models go to the GPU per call."""
    if batch < 1 or progress_every < 1:
        raise ValueError("batch and progress_every must be at least 1")
    cache_dir = Path(cache_dir)
    cache_dir.mkdir(parents=True, exist_ok=True)
    if physics is None:
        physics = make_physics()
    shots, nt_dec = physics.shots, (physics.nt - 1) // DECIMATE + 1
    memmaps = {}
    for split in splits:
        fields = np.asarray(fields_km[split], np.float32)
        n_models = len(fields)
        shape = (n_models, shots, nt_dec, N)
        bin_path = wave_path(cache_dir, split)
        done_path = cache_dir / f"waves_{split}.done.json"
        prog_path = cache_dir / f"waves_{split}.progress.json"
        if bin_path.exists() and done_path.exists():
            memmaps[split] = np.memmap(bin_path, dtype=np.float16, mode="r", shape=shape)
            continue

        start, sq_sum = 0, 0.0
        if bin_path.exists() and prog_path.exists():
            with open(prog_path, "r", encoding="utf-8") as f:
                prog = json.load(f)
            start, sq_sum = int(prog["completed"]), float(prog["sq_sum"])
        mm = np.memmap(bin_path, dtype=np.float16, mode="r+" if start > 0 else "w+", shape=shape)
        t0 = time.perf_counter()
        log(f"[waves/{split}] simulating {n_models - start} of {n_models} models on the GPU")
        pipe = _Pipeline(physics, fields, batch)
        b0 = start
        if b0 < n_models:
            pipe.enqueue(b0)
        while b0 < n_models:
            b1 = min(n_models, b0 + batch)
            if b1 < n_models:
                pipe.enqueue(b1)  # the next batch simulates while this one is written
            dec, half = pipe.take(b0)
            sums = [float(np.square(dec[k], dtype=np.float64).sum()) for k in range(b1 - b0)]
            mm[b0:b1] = half
            marks = [d for d in range(b0 + 1, b1 + 1) if d % progress_every == 0 or d == n_models]
            if marks:
                mm.flush()
            for k in range(b1 - b0):
                sq_sum += sums[k]
                done = b0 + k + 1
                if done in marks:
                    with open(prog_path, "w", encoding="utf-8") as f:
                        json.dump({"completed": done, "sq_sum": sq_sum}, f)
                    rate = (done - start) / max(1e-6, time.perf_counter() - t0)
                    log(f"[waves/{split}] {done}/{n_models} ({rate:.1f} models/s, "
                        f"ETA {(n_models - done) / max(rate, 1e-6):.0f}s)")
            b0 = b1
        rms = float(np.sqrt(sq_sum / max(1, n_models * shots * nt_dec * N)))
        with open(done_path, "w", encoding="utf-8") as f:
            json.dump({"split": split, "count": n_models, "rms": rms}, f, indent=2)
        prog_path.unlink(missing_ok=True)
        del mm
        memmaps[split] = np.memmap(bin_path, dtype=np.float16, mode="r", shape=shape)
    return memmaps


class _Pipeline:
    """Batches of decimated gathers, float32 for the square sums and float16
    for the file.  On the GPU each batch is simulated, decimated ([:, :, ::8, :])
    and cast to float16 on the current stream, then copied into pinned host
    buffers on a copy stream (two buffer sets, so batch k's copy overlaps
    batch k+1's simulation); on the CPU (a stand-in solver) it is computed
    directly."""

    def __init__(self, physics, fields: np.ndarray, batch: int):
        self.physics, self.fields, self.batch = physics, fields, batch
        self.slots: dict[int, tuple] = {}
        self.bufs: list = [None, None]
        self.copy_stream = None
        self.enqueued = 0

    def _simulate(self, b0: int):
        with torch.no_grad():
            tr = self.physics.simulate_torch(torch.from_numpy(self.fields[b0:b0 + self.batch]))
            dec = tr[:, :, ::DECIMATE, :].contiguous()
            return dec, dec.half()

    def enqueue(self, b0: int) -> None:
        dec, half = self._simulate(b0)
        if not dec.is_cuda:
            self.slots[b0] = (dec.numpy(), half.numpy(), None)
            return
        slot = self.enqueued % 2  # the other set holds the batch the host is writing
        self.enqueued += 1
        if self.bufs[slot] is None or self.bufs[slot][0].shape[0] < dec.shape[0]:
            self.bufs[slot] = (torch.empty(dec.shape, dtype=torch.float32, pin_memory=True),
                               torch.empty(dec.shape, dtype=torch.float16, pin_memory=True))
        if self.copy_stream is None:
            self.copy_stream = torch.cuda.Stream(device=dec.device)
        ready = torch.cuda.Event()
        ready.record()
        n = dec.shape[0]
        p32, p16 = self.bufs[slot][0][:n], self.bufs[slot][1][:n]
        with torch.cuda.stream(self.copy_stream):
            self.copy_stream.wait_event(ready)
            p32.copy_(dec, non_blocking=True)
            p16.copy_(half, non_blocking=True)
            copied = torch.cuda.Event()
            copied.record()
        dec.record_stream(self.copy_stream)
        half.record_stream(self.copy_stream)
        self.slots[b0] = (p32.numpy(), p16.numpy(), copied)

    def take(self, b0: int):
        dec, half, copied = self.slots.pop(b0)
        if copied is not None:
            copied.synchronize()
        return dec, half


def rms_scale(cache_dir: Path | str, fields_km: dict | None = None, physics=None) -> float:
    """Global RMS of the training gathers (data.py:181-187).  When the train
    split is not cached yet it is simulated first from fields_km["train"]."""
    done_path = Path(cache_dir) / "waves_train.done.json"
    if not done_path.exists():
        if fields_km is None:
            raise FileNotFoundError(f"{done_path} is missing and no fields were given to simulate it")
        ensure_wave_cache(fields_km, cache_dir, ("train",), physics)
    with open(done_path, "r", encoding="utf-8") as f:
        return max(float(json.load(f)["rms"]), 1e-6)


__all__ = ["SPLITS", "NT_DEC", "wave_path", "make_physics", "ensure_wave_cache", "rms_scale"]
