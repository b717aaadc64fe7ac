"""Wave-equation misfit as a PyTorch loss, batched on the GPU.

Counterpart of python/wgnet/physics.py (and the engine factory/constants of
python/wgnet/data.py:20-33, 190-193).  Same API (`lowpass_kernel`,
`WavePhysics.set_band/bandlimit/simulate/loss_and_grad`), same units (km/s
fields, gradients w.r.t. km/s) and layouts ([B, shots, nt, 70] gathers), but
every model of a batch goes to the GPU in one call instead of a Python loop
(physics.py:62-75), and `misfit()` exposes the loss as a torch.autograd.Function
whose gradient is the HIP adjoint.
"""

from __future__ import annotations

import ctypes
import math
import weakref

import numpy as np
import torch

from . import _lib
from .wave_gauss_rs import PyWaveEngine

# python/wgnet/data.py:20-29
N = 70
H = 10.0
SHOTS = 5
NT = 2001
DT = 0.0005
PEAK_HZ = 15.0
DECIMATE = 8
NT_DEC = (NT - 1) // DECIMATE + 1


def make_engine(shots: int = SHOTS, nt: int = NT, dt: float = DT, peak_hz: float = PEAK_HZ, *, nvx: int = N,
                nvz: int = N, h: float = H) -> PyWaveEngine:
    """data.py:32-33 (dt, peak_hz and the grid are parameters here; defaults
    unchanged)."""
    return PyWaveEngine(shots, nt, dt, peak_hz, 20, 1000.0, 5000.0, nvx=nvx, nvz=nvz, h=h)


def source_positions_m(shots: int = SHOTS) -> np.ndarray:
    """data.py:190-193: default source x positions in metres."""
    xmax = (N - 1) * H
    return np.array([xmax / 2.0], np.float32) if shots == 1 else np.linspace(0.0, xmax, shots, dtype=np.float32)


def lowpass_kernel(fmax_hz: float, dt: float = DT) -> np.ndarray:
    """Gaussian low-pass with its -3 dB point at fmax_hz, truncated at 4 sigma,
    unit DC gain (physics.py:18-26)."""
    sigma_s = math.sqrt(math.log(2.0) / 2.0) / (math.pi * fmax_hz)
    half = int(math.ceil(4.0 * sigma_s / dt))
    n = np.arange(-half, half + 1)
    h = np.exp(-0.5 * (n * dt / sigma_s) ** 2)
    return (h / h.sum()).astype(np.float32)


def _dev_ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _device_f32(x, device) -> torch.Tensor:
    if isinstance(x, torch.Tensor):
        return x.to(device=device, dtype=torch.float32).contiguous()
    return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(device)


def causal_filter(traces: torch.Tensor, taps: np.ndarray) -> torch.Tensor:
    """Causal FIR along time of device gathers [..., nt, nrec] with the HIP
    kernel (physics.py:48-57's conv1d, left padding len(taps) - 1)."""
    x = traces.to(dtype=torch.float32).contiguous()
    nt, nr = x.shape[-2], x.shape[-1]
    y = torch.empty_like(x)
    h = torch.from_numpy(np.ascontiguousarray(taps, np.float32)).to(x.device)
    n = x.numel() // (nt * nr)
    _lib.check(_lib.load().wg_causal_filter_dev(_dev_ptr(x), n, nt, nr, _dev_ptr(h), h.numel(), _dev_ptr(y),
                                                _stream()))
    return y


class WaveMisfit(torch.autograd.Function):
    """losses[b] = 0.5 ||p_b - d_b||^2 / ||d_b||^2 for fields in km/s.

    forward runs the HIP forward (storing snapshots), the misfit and the HIP
    adjoint in one batched call when the fields need a gradient, and the
    forward and misfit alone when they do not; backward scales the adjoint
    gradient by the upstream gradient of each model's loss."""

    @staticmethod
    def forward(ctx, fields_km: torch.Tensor, observed: torch.Tensor, physics: "WavePhysics"):
        if not ctx.needs_input_grad[0]:
            return physics.loss_torch(fields_km.detach(), observed)
        losses, grads_km = physics.loss_and_grad_torch(fields_km.detach(), observed)
        ctx.save_for_backward(grads_km)
        return losses

    @staticmethod
    def backward(ctx, g):
        (grads_km,) = ctx.saved_tensors
        return g.view(-1, 1, 1) * grads_km, None, None


class WavePhysics:
    """Owns one engine; frequency continuation by wavelet low-pass
    (physics.py:29-76)."""

    def __init__(self, shots: int = SHOTS, nt: int = NT, dt: float = DT, peak_hz: float = PEAK_HZ,
                 device: str | torch.device = "cuda", *, nvx: int = N, nvz: int = N, h: float = H) -> None:
        self.engine = make_engine(shots, nt, dt, peak_hz, nvx=nvx, nvz=nvz, h=h)
        self.shots, self.nt, self.dt = shots, nt, dt
        self.device = torch.device(device)
        self.full_wavelet = self.engine.wavelet().copy()
        self.kernel: np.ndarray | None = None
        self.fmax: float | None = None
        self._checked_obs = None  # (weakref to an observed tensor, its version) whose energies were checked

    # -- band --------------------------------------------------------------
    def set_band(self, fmax_hz: float | None) -> None:
        """None restores the full-band Ricker (physics.py:37-46)."""
        if fmax_hz is None:
            self.kernel, self.fmax = None, None
            self.engine.set_wavelet(self.full_wavelet)
            return
        self.kernel = lowpass_kernel(fmax_hz, self.dt)
        self.fmax = float(fmax_hz)
        w = np.convolve(self.full_wavelet, self.kernel, mode="full")[: self.nt].astype(np.float32)
        self.engine.set_wavelet(np.ascontiguousarray(w))

    def band_wavelet(self, fmax_hz: float | None) -> np.ndarray:
        """The source wavelet of passband fmax_hz (None: full band), as
        set_band would install it."""
        if fmax_hz is None:
            return self.full_wavelet.copy()
        k = lowpass_kernel(fmax_hz, self.dt)
        return np.convolve(self.full_wavelet, k, mode="full")[: self.nt].astype(np.float32)

    def bandlimit(self, traces):
        """Causal low-pass along time of [..., nt, nrec] data (physics.py:48-57);
        numpy in -> numpy out, tensor in -> tensor out on its device."""
        return self.bandlimit_to(traces, self.fmax)

    def bandlimit_to(self, traces, fmax_hz: float | None):
        """`bandlimit` with an explicit passband (None: unchanged)."""
        if fmax_hz is None:
            return traces
        kernel = lowpass_kernel(fmax_hz, self.dt)
        if isinstance(traces, torch.Tensor) and traces.is_cuda:
            return causal_filter(traces, kernel)
        is_np = not isinstance(traces, torch.Tensor)
        t = torch.from_numpy(np.ascontiguousarray(traces, dtype=np.float32)) if is_np else traces.float()
        lead = t.shape[:-2]
        x = t.reshape(-1, 1, t.shape[-2], t.shape[-1]).transpose(2, 3).reshape(-1, 1, t.shape[-2])
        k = torch.from_numpy(kernel).to(t.device).view(1, 1, -1)
        y = torch.nn.functional.conv1d(torch.nn.functional.pad(x, (len(kernel) - 1, 0)), k)
        out = y.view(*lead, t.shape[-1], t.shape[-2]).transpose(-1, -2).contiguous()
        return out.numpy() if is_np else out

    # -- solves --------------------------------------------------------------
    def _unit_wavelets(self, wavelets, B: int) -> torch.Tensor:
        """[B, nt] or [B, shots, nt] -> contiguous [B, shots, nt] on the device."""
        w = _device_f32(wavelets, self.device)
        if w.dim() == 2:
            w = w[:, None, :].expand(B, self.shots, self.nt)
        if tuple(w.shape) != (B, self.shots, self.nt):
            raise ValueError("wavelets must have shape [B, nt] or [B, shots, nt]")
        return w.contiguous()

    def simulate_torch(self, fields_km: torch.Tensor, wavelets=None) -> torch.Tensor:
        """[B, nvz, nvx] km/s (device) -> [B, shots, nt, nrec] (device); with
        `wavelets` ([B, nt] or [B, shots, nt]) each model uses its own."""
        vel = (_device_f32(fields_km, self.device) * 1000.0).contiguous()
        B = vel.shape[0]
        out = torch.empty((B, self.shots, self.nt, self.engine.nrec), dtype=torch.float32, device=self.device)
        lib, h = self.engine._lib, self.engine.handle
        if wavelets is None:
            _lib.check(lib.wg_simulate_dev(h, B, _dev_ptr(vel), _dev_ptr(out), _stream()))
        else:
            w = self._unit_wavelets(wavelets, B)
            _lib.check(lib.wg_simulate_wavelets_dev(h, B, _dev_ptr(vel), _dev_ptr(w), _dev_ptr(out), _stream()))
        return out

    def simulate(self, fields_km):
        """[B, 70, 70] km/s -> [B, shots, nt, 70] with the current passband
        (physics.py:59-64).  numpy in -> numpy out."""
        if isinstance(fields_km, torch.Tensor):
            return self.simulate_torch(fields_km)
        return self.simulate_torch(torch.from_numpy(np.asarray(fields_km, np.float32))).cpu().numpy()

    def illumination_torch(self, fields_km: torch.Tensor) -> torch.Tensor:
        """[B, nvz, nvx] km/s -> source illumination [B, nvz, nvx] (sum over
        shots and steps of the squared snapshot at each cell)."""
        vel = (_device_f32(fields_km, self.device) * 1000.0).contiguous()
        out = torch.empty_like(vel)
        _lib.check(self.engine._lib.wg_illumination_dev(self.engine.handle, vel.shape[0], _dev_ptr(vel),
                                                        _dev_ptr(out), _stream()))
        return out

    def _check_energy(self, observed, energy: torch.Tensor) -> None:
        """The reference asserts that the observed energy is positive
        (wave_wgpu.rs:853; NaN fails the assert too) and raises from the call.
        Reading the device energies waits for the call, so the check runs once
        per observed tensor (the same object at the same version has the same
        energies); the engine gives such models NaN losses and gradients
        either way."""
        key = self._checked_obs
        if isinstance(observed, torch.Tensor) and key is not None and key[0]() is observed \
                and key[1] == observed._version:
            return
        if bool((~(energy > 0)).any()):
            raise ValueError("Observed traces have zero energy")
        if isinstance(observed, torch.Tensor):
            self._checked_obs = (weakref.ref(observed), observed._version)

    def loss_torch(self, fields_km: torch.Tensor, observed: torch.Tensor, traces_out: torch.Tensor | None = None,
                   wavelets=None) -> torch.Tensor:
        """Per-model normalised misfit alone (no adjoint), on the device: the
        losses loss_and_grad_torch returns, bit for bit, at the cost of the
        forward."""
        vel = (_device_f32(fields_km, self.device) * 1000.0).contiguous()
        obs = _device_f32(observed, self.device)
        B = vel.shape[0]
        if obs.numel() != B * self.shots * self.nt * self.engine.nrec:
            raise ValueError("observed length does not match shots * nt * nrec")
        losses = torch.empty(B, dtype=torch.float32, device=self.device)
        energy = torch.empty(B, dtype=torch.float64, device=self.device)
        tr = _dev_ptr(traces_out) if traces_out is not None else None
        w = None if wavelets is None else self._unit_wavelets(wavelets, B)
        _lib.check(self.engine._lib.wg_loss_dev(self.engine.handle, B, _dev_ptr(vel),
                                                None if w is None else _dev_ptr(w), _dev_ptr(obs),
                                                _dev_ptr(losses), tr, _dev_ptr(energy), _stream()))
        self._check_energy(observed, energy)
        return losses

    def loss_and_grad_torch(self, fields_km: torch.Tensor, observed: torch.Tensor,
                            traces_out: torch.Tensor | None = None,
                            wavelets=None) -> tuple[torch.Tensor, torch.Tensor]:
        """Per-model normalised misfit and its gradient w.r.t. km/s, on the
        device, for a whole batch in one call; with `wavelets` ([B, nt] or
        [B, shots, nt]) each model uses its own source wavelet."""
        vel = (_device_f32(fields_km, self.device) * 1000.0).contiguous()
        obs = _device_f32(observed, self.device)
        B = vel.shape[0]
        if obs.numel() != B * self.shots * self.nt * self.engine.nrec:
            raise ValueError("observed length does not match shots * nt * nrec")
        losses = torch.empty(B, dtype=torch.float32, device=self.device)
        grads = torch.empty_like(vel)
        energy = torch.empty(B, dtype=torch.float64, device=self.device)
        tr = _dev_ptr(traces_out) if traces_out is not None else None
        lib, h = self.engine._lib, self.engine.handle
        if wavelets is None:
            _lib.check(lib.wg_loss_grad_dev(h, B, _dev_ptr(vel), _dev_ptr(obs), _dev_ptr(losses), _dev_ptr(grads), tr,
                                            _dev_ptr(energy), _stream()))
        else:
            w = self._unit_wavelets(wavelets, B)
            _lib.check(lib.wg_loss_grad_wavelets_dev(h, B, _dev_ptr(vel), _dev_ptr(w), _dev_ptr(obs),
                                                     _dev_ptr(losses), _dev_ptr(grads), tr, _dev_ptr(energy),
                                                     _stream()))
        self._check_energy(observed, energy)
        return losses, grads * 1000.0  # d/d(km/s) = 1000 d/d(m/s)

    def loss_and_grad_true_torch(self, fields_km: torch.Tensor, true_km: torch.Tensor
                                 ) -> tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
        """simulate_torch(true_km) then loss_and_grad_torch(fields_km, observed)
        in one engine call (train.py:262-263); returns (losses, grads w.r.t.
        km/s, observed [B, shots, nt, nrec]).  The observed forward runs beside
        the predicted one when the batch fits one round of the device."""
        vel = (_device_f32(fields_km, self.device) * 1000.0).contiguous()
        tvel = (_device_f32(true_km, self.device) * 1000.0).contiguous()
        B = vel.shape[0]
        if tuple(tvel.shape) != tuple(vel.shape):
            raise ValueError("true and predicted fields must have the same shape")
        obs = torch.empty((B, self.shots, self.nt, self.engine.nrec), dtype=torch.float32, device=self.device)
        losses = torch.empty(B, dtype=torch.float32, device=self.device)
        grads = torch.empty_like(vel)
        energy = torch.empty(B, dtype=torch.float64, device=self.device)
        lib, h = self.engine._lib, self.engine.handle
        _lib.check(lib.wg_loss_grad_true_dev(h, B, _dev_ptr(vel), _dev_ptr(tvel), _dev_ptr(obs), _dev_ptr(losses),
                                             _dev_ptr(grads), None, _dev_ptr(energy), _stream()))
        self._check_energy(obs, energy)
        return losses, grads * 1000.0, obs

    def loss_and_grad(self, fields_km, observed):
        """physics.py:66-76: numpy in -> (losses [B] f32, grads [B, 70, 70] f32)."""
        if isinstance(fields_km, torch.Tensor):
            return self.loss_and_grad_torch(fields_km, observed)
        losses, grads = self.loss_and_grad_torch(torch.from_numpy(np.asarray(fields_km, np.float32)),
                                                 torch.from_numpy(np.asarray(observed, np.float32)))
        return losses.cpu().numpy(), grads.cpu().numpy()

    def misfit(self, fields_km: torch.Tensor, observed: torch.Tensor) -> torch.Tensor:
        """Differentiable per-model losses [B] (torch.autograd.Function).
        Without autograd (no_grad, or fields that need no gradient) only the
        forward and the misfit run."""
        if not (torch.is_grad_enabled() and fields_km.requires_grad):
            return self.loss_torch(fields_km.detach(), observed)
        return WaveMisfit.apply(fields_km, _device_f32(observed, self.device), self)
