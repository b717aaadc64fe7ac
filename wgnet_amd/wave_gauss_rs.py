"""Drop-in replacement of the reference's PyO3 module `wave_gauss_rs`
(crates/wave-gauss-rs/src/lib.rs:10-146), backed by the HIP engine.

Same names, signatures, defaults, layouts and exception types:
  ricker_wavelet(samples, dt, peak_hz) -> float32[samples]           lib.rs:10-15
  simulate_shot_cpu(vel_mps, h, dt, nt, pml, vmin, vmax, peak_hz,
                    src_x, src_z) -> float32[nt, 70]                  lib.rs:17-42
  PyWaveEngine(shots=5, nt=2001, dt=0.0005, peak_hz=15.0, pml=20,
               vmin=1000.0, vmax=5000.0)                              lib.rs:44-138
    .shots .nt .nrec, .sources() / .set_sources(), .wavelet() /
    .set_wavelet(), .simulate(vel) -> [shots, nt, nrec],
    .evaluate_adjoint(vel, observed) -> (loss, grad [70, 70])

Differences, documented:
  - `simulate_shot_cpu` computes on the GPU with a one-shot engine; its result
    is the CPU solver's to the bit (same float32 operation order).
  - zero observed energy raises ValueError("Observed traces have zero
    energy") where the reference panics with that message (wave_wgpu.rs:853).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

N = 70  # geology-core lib.rs:5
CELLS = N * N


def _as_f32(arr, name: str, ndim: int) -> np.ndarray:
    # PyO3's PyReadonlyArray extraction: wrong type / dtype / rank -> TypeError.
    if not isinstance(arr, np.ndarray) or arr.dtype != np.float32 or arr.ndim != ndim:
        raise TypeError(f"{name}: expected a {ndim}-D numpy.ndarray of float32")
    if not arr.flags.c_contiguous:
        raise ValueError(f"{name} must be contiguous")
    return arr


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def ricker_wavelet(samples: int, dt: float, peak_hz: float) -> np.ndarray:
    """Ricker wavelet delayed by 1.5 / peak_hz (geology-core lib.rs:12-21)."""
    out = np.zeros(int(samples), np.float32)
    _lib.load().wg_ricker_wavelet(int(samples), float(dt), float(peak_hz), _ptr(out))
    return out


_SHOT_ENGINES: dict[tuple, "PyWaveEngine"] = {}


def simulate_shot_cpu(vel_mps, h, dt, nt, pml, vmin, vmax, peak_hz, src_x, src_z) -> np.ndarray:
    """Single-shot forward returning [nt, 70] traces (lib.rs:17-42; the
    solver of geology-core lib.rs:74-221)."""
    v = _as_f32(vel_mps, "vel_mps", 1)
    if v.size != CELLS:
        raise ValueError("vel_mps must have 4900 elements")
    key = (float(h), float(dt), int(nt), int(pml), float(vmin), float(vmax), float(peak_hz))
    eng = _SHOT_ENGINES.get(key)
    if eng is None:
        eng = PyWaveEngine(1, int(nt), float(dt), float(peak_hz), int(pml), float(vmin), float(vmax), h=float(h))
        _SHOT_ENGINES[key] = eng
    eng.set_sources(np.array([[src_x, src_z]], np.float32))
    return eng.simulate(v)[0]


class PyWaveEngine:
    """Multi-shot forward and exact discrete adjoint on one GPU
    (wave_wgpu.rs:95-988)."""

    def __init__(self, shots: int = 5, nt: int = 2001, dt: float = 0.0005, peak_hz: float = 15.0, pml: int = 20,
                 vmin: float = 1000.0, vmax: float = 5000.0, *, h: float = 10.0, nvx: int = N, nvz: int = N):
        lib = _lib.load()
        cfg = _lib.Config()
        lib.wg_config_default(ctypes.byref(cfg))
        cfg.shots, cfg.nt, cfg.dt, cfg.peak_hz = int(shots), int(nt), float(dt), float(peak_hz)
        cfg.pml, cfg.vmin, cfg.vmax = int(pml), float(vmin), float(vmax)
        cfg.h, cfg.nvx, cfg.nvz = float(h), int(nvx), int(nvz)
        handle = ctypes.c_void_p()
        _lib.check(lib.wg_engine_create(ctypes.byref(cfg), ctypes.byref(handle)))
        self._lib = lib
        self._h = handle
        self._shots, self._nt = int(shots), int(nt)
        self._nvx, self._nvz = int(nvx), int(nvz)
        self._nrec = lib.wg_engine_nrec(handle)
        self.dt, self.peak_hz, self.pml, self.vmin, self.vmax, self.h = float(dt), float(peak_hz), int(pml), float(vmin), float(vmax), float(h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.wg_engine_destroy(h)
            self._h = None

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    @property
    def shots(self) -> int:
        return self._shots

    @property
    def nt(self) -> int:
        return self._nt

    @property
    def nrec(self) -> int:
        return self._nrec

    @property
    def nvx(self) -> int:
        return self._nvx

    @property
    def nvz(self) -> int:
        return self._nvz

    @property
    def streamed(self) -> bool:
        """True when the grid runs on the HBM-streamed kernels."""
        return self._lib.wg_engine_streamed(self._h) == 1

    def set_snapshot_budget(self, nbytes: int) -> None:
        """Device-memory bound for snapshots and checkpoints."""
        _lib.check(self._lib.wg_set_snapshot_budget(self._h, int(nbytes)))

    def set_time_block(self, steps: int) -> None:
        """Adjoint time steps per HBM-streamed launch (1, or 2, the default);
        results are identical either way."""
        _lib.check(self._lib.wg_set_time_block(self._h, int(steps)))

    def set_forward_time_block(self, steps: int) -> None:
        """Forward time steps per HBM-streamed launch (1, or 2, the default);
        results are identical either way."""
        _lib.check(self._lib.wg_set_forward_time_block(self._h, int(steps)))

    @property
    def bands(self) -> int:
        """Row bands per (model, shot) unit on the band kernels (0: none)."""
        return int(self._lib.wg_engine_bands(self._h))

    def set_band_split(self, on) -> None:
        """Run a grid that splits into row bands on the band kernels (True or
        1, the default: compiled for the grid where it has them; 2: the
        runtime-geometry band kernels) or on the HBM-streamed kernels (False
        or 0); results are identical."""
        _lib.check(self._lib.wg_set_band_split(self._h, int(on)))

    def set_stream_tile(self, tile: int) -> None:
        """Tile shape of the HBM-streamed kernels (0 .. wg_stream_tiles() - 1)."""
        _lib.check(self._lib.wg_set_stream_tile(self._h, int(tile)))

    def dims(self) -> tuple[int, int, int]:
        nx, nz, pad = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        _lib.check(self._lib.wg_engine_dims(self._h, ctypes.byref(nx), ctypes.byref(nz), ctypes.byref(pad)))
        return nx.value, nz.value, pad.value

    def sources(self) -> np.ndarray:
        """Current source positions [shots, 2] in metres (lib.rs:74-78)."""
        out = np.zeros((self._shots, 2), np.float32)
        _lib.check(self._lib.wg_get_sources(self._h, _ptr(out), self._shots))
        return out

    def set_sources(self, sources) -> None:
        """Update source positions [shots, 2] in metres (lib.rs:80-90)."""
        if not isinstance(sources, np.ndarray) or sources.dtype != np.float32 or sources.ndim != 2:
            raise TypeError("sources: expected a 2-D numpy.ndarray of float32")
        if sources.shape[0] != self._shots or sources.shape[1] != 2:
            raise ValueError("sources must have shape [shots, 2]")
        if not sources.flags.c_contiguous:
            raise ValueError("sources must be contiguous")
        _lib.check(self._lib.wg_set_sources(self._h, _ptr(sources), self._shots))

    def wavelet(self) -> np.ndarray:
        """Current source wavelet [nt] (lib.rs:92-95)."""
        out = np.zeros(self._nt, np.float32)
        _lib.check(self._lib.wg_get_wavelet(self._h, _ptr(out), self._nt))
        return out

    def set_wavelet(self, wavelet) -> None:
        """Replace the source wavelet [nt] (lib.rs:97-106)."""
        w = _as_f32(wavelet, "wavelet", 1)
        if w.size != self._nt:
            raise ValueError("wavelet length must equal nt")
        _lib.check(self._lib.wg_set_wavelet(self._h, _ptr(w), self._nt))

    def simulate(self, vel_mps) -> np.ndarray:
        """[nvz*nvx] m/s -> [shots, nt, nrec] float32 (lib.rs:108-116)."""
        v = _as_f32(vel_mps, "vel_mps", 1)
        if v.size != self._nvx * self._nvz:
            raise ValueError(f"vel_mps must have {self._nvx * self._nvz} elements")
        out = np.zeros((self._shots, self._nt, self._nrec), np.float32)
        _lib.check(self._lib.wg_simulate(self._h, _ptr(v), _ptr(out)))
        return out

    def evaluate_adjoint(self, vel_mps, observed) -> tuple[float, np.ndarray]:
        """(loss, grad [nvz, nvx]) with loss = 0.5 ||p - d||^2 / ||d||^2 and
        grad = dloss / d(v in m/s) (lib.rs:118-137)."""
        v = _as_f32(vel_mps, "vel_mps", 1)
        o = _as_f32(observed, "observed", 1)
        if v.size != self._nvx * self._nvz:
            raise ValueError(f"vel_mps must have {self._nvx * self._nvz} elements")
        if o.size != self._shots * self._nt * self._nrec:
            raise ValueError("observed length does not match shots * nt * nrec")
        loss = np.zeros(1, np.float32)
        grad = np.zeros((self._nvz, self._nvx), np.float32)
        _lib.check(self._lib.wg_evaluate_adjoint(self._h, _ptr(v), _ptr(o), _ptr(loss), _ptr(grad)))
        return float(loss[0]), grad


def install() -> None:
    """Registers this module as `wave_gauss_rs`, so the reference's
    `import wave_gauss_rs` (python/wgnet/data.py:18) resolves to it."""
    import sys

    sys.modules["wave_gauss_rs"] = sys.modules[__name__]
