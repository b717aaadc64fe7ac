"""Pixel-grid FWI (fwi_eval.py:29-63) on the CPU with a stand-in physics:
the lockstep batch gives every problem exactly the trajectory `fwi` gives it
alone (final model and per-band history, bit for bit), problems that finish
early leave the batch, a failed batched call reaches every problem without a
hang, bounds hold, and the macro-model start is the mirror-padded separable
Gaussian (igm.py:52-63) that scipy.ndimage computes."""

from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.ndimage as nd
import torch

from wgnet_amd import fwi_eval

N = 70


class _StandIn:
    """gathers(f) = w(band) * (f + 0.1 f^2) on a [1, 70, 70] 'recording';
    misfit 0.5 ||g(f) - d||^2 and its gradient, in float32, one model at a
    time whatever the batch."""

    device = torch.device("cpu")

    def __init__(self, fail_after=None):
        self.fmax = None
        self.fail_after, self.batched_calls = fail_after, 0

    @staticmethod
    def factor(band):
        return np.float32(1.0 if band is None else 0.5 + band / 20.0)

    def set_band(self, band):
        self.fmax = band

    def band_wavelet(self, band):
        return np.full(8, self.factor(band), np.float32)

    def bandlimit_to(self, x, band):
        return (np.asarray(x, np.float32) * self.factor(band)).astype(np.float32)

    def bandlimit(self, x):
        return self.bandlimit_to(x, self.fmax)

    @staticmethod
    def _one(f, d, w):
        r = w * (f + 0.1 * f * f)[None] - d
        return 0.5 * (r * r).sum(), w * r[0] * (1 + 0.2 * f)

    def loss_and_grad(self, fields, observed):
        f, d = torch.from_numpy(np.asarray(fields, np.float32)), torch.from_numpy(np.asarray(observed, np.float32))
        out = [self._one(f[b], d[b], torch.tensor(self.factor(self.fmax))) for b in range(f.shape[0])]
        return (torch.stack([o[0] for o in out]).numpy(), torch.stack([o[1] for o in out]).numpy())

    def loss_and_grad_torch(self, fields, observed, wavelets=None):
        self.batched_calls += 1
        if self.fail_after is not None and self.batched_calls > self.fail_after:
            raise RuntimeError("device failure")
        out = [self._one(fields[b], observed[b], wavelets[b, 0]) for b in range(fields.shape[0])]
        return torch.stack([o[0] for o in out]), torch.stack([o[1] for o in out])


def _problems(P, seed=0):
    rng = np.random.default_rng(seed)
    truth = (2.0 + 2.0 * rng.random((P, N, N))).astype(np.float32)
    observed = (truth + np.float32(0.1) * truth * truth)[:, None].astype(np.float32)  # the full-band recording
    starts = np.stack([np.full((N, N), 1.6 + 0.5 * p, np.float32) for p in range(P)])
    return starts, observed


@pytest.mark.timeout(120)
@pytest.mark.parametrize("workers", [0, 2, None])
def test_lockstep_batch_equals_one_problem_at_a_time(workers):
    """In-process threads (0), two worker processes, and the default."""
    bands, budget = [5.0, 12.0, None], 18
    starts, observed = _problems(4)
    phys = _StandIn()
    finals, hists, calls = fwi_eval.fwi_batch(starts, observed, phys, bands, budget, workers=workers)
    solves = []
    for p in range(4):
        f1, h1 = fwi_eval.fwi(starts[p], observed[p], _StandIn(), bands, budget)
        assert f1.tobytes() == finals[p].tobytes()
        assert h1 == hists[p]
        solves.append(sum(h["solves"] for h in h1))
        assert f1.min() >= fwi_eval.V_BOUNDS_KM[0] and f1.max() <= fwi_eval.V_BOUNDS_KM[1]
        ph = _StandIn()
        assert fwi_eval.misfit_full(f1, observed[p], ph) < fwi_eval.misfit_full(starts[p], observed[p], ph)
    # one call per round: at least the longest problem's solves, at most their sum
    assert max(solves) <= calls <= sum(solves) and calls == phys.batched_calls


@pytest.mark.timeout(120)
@pytest.mark.parametrize("workers", [0, 2])
def test_problems_finishing_early_leave_the_batch(workers):
    """Problems whose start is already the answer stop after a few solves;
    the others keep running in smaller batches."""
    bands, budget = [None], 30
    starts, observed = _problems(3, seed=1)
    rng = np.random.default_rng(1)
    truth = (2.0 + 2.0 * rng.random((3, N, N))).astype(np.float32)  # as _problems(3, seed=1) draws it
    starts[1] = truth[1]  # converged at once
    phys = _StandIn()
    finals, hists, calls = fwi_eval.fwi_batch(starts, observed, phys, bands, budget, workers=workers)
    s = [h[0]["solves"] for h in hists]
    assert s[1] < min(s[0], s[2])
    assert calls == max(s)
    for p in range(3):
        f1, h1 = fwi_eval.fwi(starts[p], observed[p], _StandIn(), bands, budget)
        assert f1.tobytes() == finals[p].tobytes() and h1 == hists[p]


@pytest.mark.timeout(60)
@pytest.mark.parametrize("workers", [0, 2])
def test_failed_batched_call_reaches_every_problem(workers):
    starts, observed = _problems(3)
    with pytest.raises(RuntimeError):
        fwi_eval.fwi_batch(starts, observed, _StandIn(fail_after=2), [None], 10, workers=workers)


def test_fwi_restores_full_band():
    starts, observed = _problems(1)
    phys = _StandIn()
    fwi_eval.fwi(starts[0], observed[0], phys, [5.0, 8.0], 4)
    assert phys.fmax is None
    # maxfun = maxiter = max(2, budget // bands) per band
    _, h = fwi_eval.fwi(starts[0], observed[0], phys, [5.0, 8.0, 12.0, None], 4)
    assert all(1 <= x["solves"] <= 2 + 2 for x in h)


@pytest.mark.parametrize("sigma", [1.5, 6.0])
def test_macro_start_is_mirror_padded_gaussian(sigma):
    rng = np.random.default_rng(3)
    v = (1.5 + 3.0 * rng.random((2, N, N))).astype(np.float32)
    half = math.ceil(3.0 * sigma)
    t = np.arange(-half, half + 1)
    w = np.exp(-0.5 * (t / sigma) ** 2)
    w /= w.sum()
    ref = nd.correlate1d(nd.correlate1d(v.astype(np.float64), w, axis=2, mode="mirror"), w, axis=1, mode="mirror")
    m = fwi_eval.macro_start(v, sigma)
    assert m.dtype == np.float32 and m.shape == v.shape
    assert np.abs(m - ref).max() <= 1e-5
    assert np.abs(fwi_eval.macro_start(np.full((1, N, N), 2.5, np.float32), sigma) - 2.5).max() <= 1e-6
    assert np.array_equal(fwi_eval.macro_start(v, 0.0), v)


@pytest.mark.timeout(60)
def test_worker_failure_is_reported():
    """An exception inside a worker (here: an empty band list) comes back as
    a RuntimeError carrying the worker's traceback, and no worker is left."""
    starts, observed = _problems(2)
    with pytest.raises(RuntimeError, match="ZeroDivisionError"):
        fwi_eval.fwi_batch(starts, observed, _StandIn(), [], 6, workers=2)
