"""numpy restatement of the reference's low-pass and causal band-limit
(python/wgnet/physics.py:18-57).  TEST INFRASTRUCTURE ONLY."""

from __future__ import annotations

import math

import numpy as np


def lowpass_kernel(fmax_hz: float, dt: float) -> np.ndarray:
    """physics.py:18-26."""
    sigma_s = math.sqrt(math.log(2.0) / 2.0) / (math.pi * fmax_hz)
    half = int(math.ceil(4.0 * sigma_s / dt))
    n = np.arange(-half, half + 1)
    h = np.exp(-0.5 * (n * dt / sigma_s) ** 2)
    return (h / h.sum()).astype(np.float32)


def causal_filter(traces: np.ndarray, kernel: np.ndarray) -> np.ndarray:
    """Causal convolution along axis -2 truncated to the input length: the
    result of physics.py:48-57 (conv1d on left-zero-padded traces)."""
    t = np.asarray(traces, np.float64)
    k = np.asarray(kernel, np.float64)
    nt = t.shape[-2]
    out = np.zeros_like(t)
    for j in range(len(k)):
        out[..., j:, :] += k[j] * t[..., : nt - j, :] if j < nt else 0.0
    return out
