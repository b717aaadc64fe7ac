"""Uneven batches: many solves with different grids, step counts, shot counts,
source positions and passbands, run as one HIP graph.

No reference counterpart: the reference solves one 70 x 70 model per engine
call with the engine's own source list (python/wgnet/physics.py:62-75,
crates/wave-gauss-rs/src/wave_wgpu.rs:761-988).  Here a batch of solves is
grouped by (grid, nt); each group is one engine whose units are (solve, shot
slot) with up to `max_shots` slots per solve.  Every unit carries its own
source position and its own band-limited wavelet; slots beyond a solve's shot
count are inactive padding (zero traces, zero gradient, no time steps).

By default (`fused=True`) the whole batch is one engine call
(wg_loss_grad_groups_dev): every group's units (70 x 70) or row bands
(100 x 100, 141 x 141) are the tasks of ONE forward launch and ONE adjoint
launch, longest first, with each group's energy, residual, loss and fold
around them.  With `fused=False` (or a grid without compiled kernels) each
group is its own engine call: one launch of the resident or band kernels, or
the HBM-streamed per-step kernels with the tile shape and adjoint steps per
launch that the autotuner picked for that grid, the groups side by side on a
few streams (longest first onto the least loaded).  Either way the batch is
captured once as a HIP graph and replayed, and the results are the same
bits.

Each solve applies the causal low-pass of its passband (physics.py:18-57) to
both its wavelet and its observed data.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .physics import causal_filter, lowpass_kernel
from .wave_gauss_rs import PyWaveEngine


@dataclass
class Solve:
    """One solve: a velocity model on an nvx x nvz grid, nt steps, sources at
    x positions src_x (metres, depth src_z), passband band_hz (None: full)."""

    nvx: int
    nvz: int
    nt: int
    src_x: np.ndarray
    band_hz: float | None = None
    src_z: float = 10.0

    @property
    def shots(self) -> int:
        return int(len(self.src_x))


def _p(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


@dataclass
class _Group:
    key: tuple
    ids: list
    engine: PyWaveEngine = None
    cost: float = 0.0
    lane: int = 0
    buf: dict = field(default_factory=dict)


class UnevenBatch:
    """Fixed batch geometry; models and observed data are set per call."""

    def __init__(self, solves: list[Solve], dt: float = 0.001, peak_hz: float = 15.0, h: float = 10.0,
                 max_shots: int = 5, memory_fraction: float = 0.6, lanes: int | None = None, fused: bool = True):
        self.solves = list(solves)
        self.fused = bool(fused)
        self.dt, self.peak_hz, self.h, self.max_shots = float(dt), float(peak_hz), float(h), int(max_shots)
        self._lib = _lib.load()
        self.graph: torch.cuda.CUDAGraph | None = None
        self.stream = torch.cuda.Stream()
        groups: dict = {}
        for i, s in enumerate(self.solves):
            if not 1 <= s.shots <= self.max_shots:
                raise ValueError(f"solve {i}: shots must be in [1, {self.max_shots}]")
            groups.setdefault((s.nvx, s.nvz, s.nt), []).append(i)
        self.groups = [_Group(k, ids) for k, ids in groups.items()]
        free, _ = torch.cuda.mem_get_info()
        need = []
        for g in self.groups:
            nvx, nvz, nt = g.key
            inner = (nvx + 40) * (nvz + 40)  # (nx - 4) (nz - 4) with the 22-cell pad
            g.cost = float(inner) * nt * sum(self.solves[i].shots for i in g.ids)
            need.append(float(inner) * nt * 4 * self.max_shots * len(g.ids))
        total_need = sum(need)
        budget = memory_fraction * free
        for g, nd in zip(self.groups, need):
            self._setup_group(g, max(int(budget * nd / max(total_need, 1.0)), 1 << 26))
        self.groups.sort(key=lambda g: -g.cost)  # longest first
        # Groups are independent (an engine and buffers each), so they run on
        # `lanes` streams side by side (default: one per group, at most 16):
        # the small grids' per-step launches are latency-bound and leave the
        # device half idle one after another.  Longest first onto the least
        # loaded lane.
        n_lanes = min(len(self.groups), 16) if lanes is None else min(int(lanes), len(self.groups))
        self.lanes = [torch.cuda.Stream() for _ in range(max(1, n_lanes))]
        load = [0.0] * len(self.lanes)
        for g in self.groups:
            g.lane = min(range(len(load)), key=load.__getitem__)
            load[g.lane] += g.cost
        # the arguments of the fused call (every buffer is allocated once)
        self._group_args = (_lib.GroupArgs * len(self.groups))()
        for k, g in enumerate(self.groups):
            b = g.buf
            self._group_args[k] = _lib.GroupArgs(
                g.engine.handle, len(g.ids), b["cell"].data_ptr(), b["w"].data_ptr(), b["active"].data_ptr(),
                b["vel"].data_ptr(), b["wav"].data_ptr(), b["obs"].data_ptr(), b["loss"].data_ptr(),
                b["grad"].data_ptr(), b["traces"].data_ptr(), b["energy"].data_ptr())

    # -- set-up --------------------------------------------------------------
    def _setup_group(self, g: _Group, snap_budget: int) -> None:
        nvx, nvz, nt = g.key
        S, B = self.max_shots, len(g.ids)
        g.engine = PyWaveEngine(S, nt, self.dt, self.peak_hz, 20, 1000.0, 5000.0, h=self.h, nvx=nvx, nvz=nvz)
        g.engine.set_snapshot_budget(snap_budget)
        lib, eh = self._lib, g.engine.handle
        xz = np.zeros((B * S, 2), np.float32)
        active = np.zeros(B * S, np.uint8)
        full = g.engine.wavelet()
        wav = np.zeros((B, S, nt), np.float32)
        wav_full = np.zeros((B, S, nt), np.float32)
        for b, i in enumerate(g.ids):
            s = self.solves[i]
            w = full if s.band_hz is None else np.convolve(full, lowpass_kernel(s.band_hz, self.dt),
                                                           mode="full")[:nt].astype(np.float32)
            for k in range(S):
                kk = min(k, s.shots - 1)  # padding slots repeat a valid position
                xz[b * S + k] = (s.src_x[kk], s.src_z)
                if k < s.shots:
                    active[b * S + k] = 1
                    wav[b, k] = w
                    wav_full[b, k] = full
        cell = np.zeros(B * S * 4, np.int32)
        wgt = np.zeros(B * S * 4, np.float32)
        _lib.check(lib.wg_source_corners(eh, B * S, ctypes.c_void_p(xz.ctypes.data), ctypes.c_void_p(cell.ctypes.data),
                                         ctypes.c_void_p(wgt.ctypes.data)))
        R = g.engine.nrec
        dev = "cuda"
        g.buf = {
            "cell": torch.from_numpy(cell).to(dev),
            "w": torch.from_numpy(wgt).to(dev),
            "active": torch.from_numpy(active).to(dev),
            "wav": torch.from_numpy(wav).to(dev),
            "wav_full": torch.from_numpy(wav_full).to(dev),
            "vel": torch.zeros((B, nvz, nvx), device=dev),
            "obs": torch.zeros((B, S, nt, R), device=dev),
            "loss": torch.zeros(B, device=dev),
            "grad": torch.zeros((B, nvz, nvx), device=dev),
            "energy": torch.zeros(B, dtype=torch.float64, device=dev),
            "traces": torch.zeros((B, S, nt, R), device=dev),
        }

    @property
    def active_shot_solves(self) -> int:
        return sum(s.shots for s in self.solves)

    def autotune(self, repeats: int = 2) -> dict:
        """For each streamed group: times every tile shape of the HBM-streamed
        kernels on the group's own forward and keeps the fastest, then times
        the group's loss and gradient with one and two adjoint steps per
        launch and keeps the faster; returns {(nvx, nvz, nt): (tile, steps)}.
        Results do not depend on either choice."""
        lib = self._lib

        def timed(fn) -> float:
            fn()  # warm-up
            ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ev0.record()
            for _ in range(repeats):
                fn()
            ev1.record()
            torch.cuda.synchronize()
            return ev0.elapsed_time(ev1)

        chosen = {}
        for g in self.groups:
            if not g.engine.streamed or lib.wg_engine_bands(g.engine.handle) > 0:
                continue  # resident kernels: whole units or row bands
            best, best_t = 0, math.inf
            for tile in range(lib.wg_stream_tiles()):
                _lib.check(lib.wg_set_stream_tile(g.engine.handle, tile))
                t = timed(lambda: self._simulate_group(g))
                if t < best_t:
                    best, best_t = tile, t
            _lib.check(lib.wg_set_stream_tile(g.engine.handle, best))
            steps, steps_t = 1, math.inf
            for tb in (1, 2):
                _lib.check(lib.wg_set_time_block(g.engine.handle, tb))
                t = timed(lambda: self._run_group(g))
                if t < steps_t:
                    steps, steps_t = tb, t
            _lib.check(lib.wg_set_time_block(g.engine.handle, steps))
            chosen[g.key] = (best, steps)
        self.graph = None
        return chosen

    # -- data ----------------------------------------------------------------
    def set_models(self, vel_mps: list) -> None:
        """Velocity model [nvz, nvx] (m/s) of every solve, in solve order."""
        for g in self.groups:
            for b, i in enumerate(g.ids):
                g.buf["vel"][b].copy_(_dev(vel_mps[i]))

    def set_observed(self, observed: list) -> None:
        """Observed gathers [shots, nt, nvx] of every solve (full band); each is
        low-passed to its solve's passband."""
        for g in self.groups:
            obs = g.buf["obs"]
            obs.zero_()
            for b, i in enumerate(g.ids):
                s = self.solves[i]
                d = _dev(observed[i])
                if s.band_hz is not None:
                    d = _bandlimit(d, s.band_hz, self.dt)
                obs[b, : s.shots].copy_(d)

    def _simulate_group(self, g: _Group, full_band: bool = False) -> torch.Tensor:
        B = len(g.ids)
        b = g.buf
        wav = b["wav_full"] if full_band else b["wav"]
        _lib.check(self._lib.wg_simulate_units_dev(g.engine.handle, B, _p(b["cell"]), _p(b["w"]), _p(b["active"]),
                                                   _p(b["vel"]), _p(wav), _p(b["traces"]), _stream()))
        return b["traces"]

    def simulate(self, full_band: bool = False) -> list:
        """Gathers [shots, nt, nvx] of the current models, band-limited per
        solve (or with the full-band wavelet)."""
        out = [None] * len(self.solves)
        for g in self.groups:
            tr = self._simulate_group(g, full_band)
            for b, i in enumerate(g.ids):
                out[i] = tr[b, : self.solves[i].shots].clone()
        return out

    # -- loss and gradient ---------------------------------------------------
    def _run_group(self, g: _Group) -> None:
        b = g.buf
        _lib.check(self._lib.wg_loss_grad_units_dev(
            g.engine.handle, len(g.ids), _p(b["cell"]), _p(b["w"]), _p(b["active"]), _p(b["vel"]), _p(b["wav"]),
            _p(b["obs"]), _p(b["loss"]), _p(b["grad"]), _p(b["traces"]), _p(b["energy"]), _stream()))

    def _run(self) -> None:
        """Fused: one engine call for every group on the current stream.
        Else every group on its lane, forked from and joined back into the
        current stream (also inside a graph capture)."""
        if self.fused:
            _lib.check(self._lib.wg_loss_grad_groups_dev(len(self.groups), ctypes.byref(self._group_args), _stream()))
            return
        cur = torch.cuda.current_stream()
        fork = torch.cuda.Event()
        fork.record(cur)
        joins = []
        for k, lane in enumerate(self.lanes):
            lane.wait_event(fork)
            with torch.cuda.stream(lane):
                for g in self.groups:
                    if g.lane == k:
                        self._run_group(g)
            e = torch.cuda.Event()
            e.record(lane)
            joins.append(e)
        for e in joins:
            cur.wait_event(e)

    def run_eager(self) -> None:
        """Loss and gradient of every solve, launched directly."""
        self._run()

    def capture(self) -> None:
        """Captures the whole batch as one HIP graph (after a warm-up run on
        the capture stream that allocates every workspace)."""
        torch.cuda.synchronize()
        with torch.cuda.stream(self.stream):
            self._run()
        self.stream.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=self.stream):
            self._run()
        self.graph = g

    def replay(self) -> None:
        if self.graph is None:
            self.capture()
        self.graph.replay()

    def results(self) -> list:
        """(loss, grad [nvz, nvx] in m/s units) per solve, as device tensors."""
        out = [None] * len(self.solves)
        for g in self.groups:
            for b, i in enumerate(g.ids):
                out[i] = (g.buf["loss"][b], g.buf["grad"][b])
        return out

    def energies(self) -> list:
        out = [None] * len(self.solves)
        for g in self.groups:
            for b, i in enumerate(g.ids):
                out[i] = g.buf["energy"][b]
        return out


def _dev(x) -> torch.Tensor:
    """numpy array or tensor -> float32 tensor on the device."""
    if isinstance(x, torch.Tensor):
        return x.to(device="cuda", dtype=torch.float32)
    return torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda()


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _bandlimit(d: torch.Tensor, fmax_hz: float, dt: float) -> torch.Tensor:
    """Causal low-pass along time of device gathers [..., nt, nrec]
    (physics.py:48-57), HIP kernel."""
    return causal_filter(d, lowpass_kernel(fmax_hz, dt))


def random_solves(n: int, seed: int, grids=(70, 100, 141), nts=(1000, 1500, 2000), bands=(6.0, 9.0, 12.0, 15.0),
                  max_shots: int = 5, h: float = 10.0) -> list[Solve]:
    """n solves with grid, nt, shot count (1..max_shots), source x positions
    (uniform over the surface) and passband drawn uniformly."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        g = int(rng.choice(grids))
        nt = int(rng.choice(nts))
        shots = int(rng.integers(1, max_shots + 1))
        xs = np.sort(rng.uniform(0.0, (g - 1) * h, shots)).astype(np.float32)
        out.append(Solve(g, g, nt, xs, float(rng.choice(bands))))
    return out
