"""Synthetic velocity models for the benchmark and tests (OpenFWI data is not
available here).  Velocities in m/s on an [nvz, nvx] grid, row-major [z, x].

Families:
  flat     n layers, interface depths uniform in [10, 60] cells (scaled to
           the grid), velocities uniform in [1500, 4500] sorted with depth
  curved   the same with sinusoidal interfaces: amplitude U[0, 8] cells,
           wavelength U[20, 70] cells
  faulted  flat layers cut by one dipping fault with a throw of U{2..10} cells
  anomaly  isotropic Gaussian, sigma U[3, 8] cells, amplitude +-U[0, 300] m/s
  large    1000 x 300: 8 curved layers plus 20 Gaussian anomalies (sigma
           U[5, 30] cells, amplitude +-U[0, 400] m/s), clipped to [1500, 4500]
"""

from __future__ import annotations

import numpy as np


def _velocities(rng, n):
    return np.sort(rng.uniform(1500.0, 4500.0, n))


def _interfaces(rng, n_if, nvz):
    lo, hi = max(1, round(10 * nvz / 70)), max(2, round(60 * nvz / 70))
    return np.sort(rng.integers(lo, hi + 1, n_if)).astype(np.float64)


def _fill(depth_of_interface, vels, nvz, nvx):
    """depth_of_interface: [n_if, nvx] depth (cells) of each interface per column."""
    z = np.arange(nvz, dtype=np.float64)[:, None]
    m = np.full((nvz, nvx), vels[0])
    for k in range(depth_of_interface.shape[0]):
        m = np.where(z >= depth_of_interface[k][None, :], vels[k + 1], m)
    return m


def flat(rng, nvz=70, nvx=70, layers=3):
    d = _interfaces(rng, layers - 1, nvz)
    return _fill(np.repeat(d[:, None], nvx, 1), _velocities(rng, layers), nvz, nvx).astype(np.float32)


def curved(rng, nvz=70, nvx=70, layers=3):
    d = _interfaces(rng, layers - 1, nvz)
    x = np.arange(nvx, dtype=np.float64)
    rows = []
    for k in range(layers - 1):
        amp, lam, ph = rng.uniform(0, 8), rng.uniform(20, 70), rng.uniform(0, 2 * np.pi)
        rows.append(d[k] + amp * np.sin(2 * np.pi * x / lam + ph))
    return _fill(np.array(rows), _velocities(rng, layers), nvz, nvx).astype(np.float32)


def faulted(rng, nvz=70, nvx=70, layers=3):
    d = _interfaces(rng, layers - 1, nvz)
    throw = int(rng.integers(2, 11))
    x0 = rng.uniform(0.25 * nvx, 0.75 * nvx)
    dip = rng.uniform(0.3, 1.5) * rng.choice([-1.0, 1.0])  # dx per dz of the fault trace
    z = np.arange(nvz, dtype=np.float64)[:, None]
    x = np.arange(nvx, dtype=np.float64)[None, :]
    hanging = x > x0 + dip * z  # one side of the fault plane is dropped by `throw`
    vels = _velocities(rng, layers)
    m = np.full((nvz, nvx), vels[0])
    for k in range(layers - 1):
        depth = np.where(hanging, d[k] + throw, d[k])
        m = np.where(z >= depth, vels[k + 1], m)
    return m.astype(np.float32)


def gaussian_anomaly(rng, nvz=70, nvx=70, sigma=(3.0, 8.0), amp=300.0):
    s = rng.uniform(*sigma)
    a = rng.uniform(0.0, amp) * rng.choice([-1.0, 1.0])
    cz, cx = rng.uniform(0.15 * nvz, 0.85 * nvz), rng.uniform(0.15 * nvx, 0.85 * nvx)
    z, x = np.mgrid[0:nvz, 0:nvx]
    return (a * np.exp(-((x - cx) ** 2 + (z - cz) ** 2) / (2 * s * s))).astype(np.float32)


FAMILIES = (flat, curved, faulted)


def batch(n: int, seed: int, nvz: int = 70, nvx: int = 70):
    """(start, true) [n, nvz, nvz] m/s: family b % 3; true = start + anomaly,
    clipped to [1500, 4500]."""
    start = np.empty((n, nvz, nvx), np.float32)
    true = np.empty_like(start)
    for b in range(n):
        rng = np.random.default_rng(seed + b)
        m = FAMILIES[b % 3](rng, nvz, nvx)
        start[b] = m
        true[b] = np.clip(m + gaussian_anomaly(rng, nvz, nvx), 1500.0, 4500.0)
    return start, true


def large_batch(n: int, seed: int, nvz: int = 300, nvx: int = 1000, layers: int = 8, anomalies: int = 20):
    """(start, true) [n, nvz, nvx] m/s for the streamed large-grid workload:
    start = curved layers, true = start + anomalies, both clipped to [1500, 4500]."""
    start = np.empty((n, nvz, nvx), np.float32)
    true = np.empty_like(start)
    for b in range(n):
        rng = np.random.default_rng(seed + b)
        m = curved(rng, nvz, nvx, layers)
        a = np.zeros((nvz, nvx), np.float32)
        for _ in range(anomalies):
            a += gaussian_anomaly(rng, nvz, nvx, sigma=(5.0, 30.0), amp=400.0)
        start[b] = np.clip(m, 1500.0, 4500.0)
        true[b] = np.clip(m + a, 1500.0, 4500.0)
    return start, true
