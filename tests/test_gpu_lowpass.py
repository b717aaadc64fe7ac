"""HIP causal low-pass (physics.py:48-57) against the numpy oracle and
against torch's conv1d on the CPU in float64 (the reference's own call)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("fmax,dt", [(6.0, 0.001), (9.0, 0.001), (12.0, 0.0008), (15.0, 0.0005)])
def test_causal_filter_matches_oracle(fmax, dt):
    from oracle.bandlimit import causal_filter as ref_filter
    from oracle.bandlimit import lowpass_kernel as ref_kernel
    from wgnet_amd.physics import causal_filter, lowpass_kernel

    k = lowpass_kernel(fmax, dt)
    assert np.array_equal(k, ref_kernel(fmax, dt))
    rng = np.random.default_rng(int(fmax))
    x = rng.standard_normal((3, 700, 70)).astype(np.float32)
    y = causal_filter(torch.from_numpy(x).cuda(), k).cpu().numpy()
    err = relerr(y, ref_filter(x, k))
    print(f"causal filter {fmax} Hz, dt {dt}: {len(k)} taps, rel-L2 {err:.2e}")
    assert err <= 1e-6


def test_causal_filter_is_conv1d_with_left_padding():
    """A non-symmetric tap vector pins the orientation: conv1d cross-correlates."""
    from wgnet_amd.physics import causal_filter

    rng = np.random.default_rng(9)
    h = rng.standard_normal(37).astype(np.float32)
    x = rng.standard_normal((2, 200, 5)).astype(np.float32)
    y = causal_filter(torch.from_numpy(x).cuda(), h).cpu().numpy()
    xt = torch.from_numpy(x.astype(np.float64)).permute(0, 2, 1).reshape(-1, 1, 200)
    ref = torch.nn.functional.conv1d(torch.nn.functional.pad(xt, (36, 0)), torch.from_numpy(h.astype(np.float64)).view(1, 1, -1))
    ref = ref.view(2, 5, 200).permute(0, 2, 1).numpy()
    assert relerr(y, ref) <= 1e-6
