"""Learned gradient preconditioner for refine(): the HIP network of
csrc/precond.hip behind a small Python class, its input features, the
training-pair generator and the refine hook.

No reference counterpart: the reference refines with plain projected L-BFGS
(python/wgnet/refine.py:92-169).  The network sees (velocity-space gradient,
normalised velocity) and predicts the model update (true - current), the pairs
coming from the batched wave solver itself.  In refine() it is applied at the
gradient hook of refine.py:105-106 in one of two ways:
  update_hook()  the predicted update itself (km/s), which refine_batch fits
                 onto the scene parameters and tries as the first step of
                 each band (`update=`);
  hook()         minus the predicted update rescaled to the gradient's norm,
                 a gradient-like direction (`precondition=`, `pc_mode`).
Plain L-BFGS steps remain the fallback whenever a learned step fails its
line search.
"""

from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

VEL_CENTRE_KM = 3.0
VEL_SCALE_KM = 1.5
TARGET_SCALE_KM = 0.3
ASINH_GAIN = 3.0  # compressed inputs: asinh(3 g / median RMS)


def _p(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def _stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def features(grads: torch.Tensor, fields_km: torch.Tensor, grad_scale: float, compress: bool = False) -> torch.Tensor:
    """[N, H, W] gradient and field (km/s) -> network input [N, 2, H, W]:
    grad_scale * gradient (one constant for all models, so the input keeps
    the gradient's magnitude, which grows with the size of the model error;
    asinh-compressed when `compress`), and (v - 3) / 1.5."""
    g = grads.float() * grad_scale
    if compress:
        g = torch.asinh(g)
    return torch.stack([g, (fields_km.float() - VEL_CENTRE_KM) / VEL_SCALE_KM], 1).contiguous()


class Preconditioner:
    """bf16 network of dilated 3x3 convolutions 2 -> 32 -> 32 -> 32 -> 1
    (dilations 4, 16, 8, 1; GELU) on MFMA."""

    def __init__(self, max_batch: int = 256, height: int = 70, width: int = 70, seed: int = 0):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        _lib.check(self._lib.wg_pc_create(int(max_batch), int(height), int(width), int(seed), ctypes.byref(h)), pc=True)
        self._h = h
        self.max_batch, self.height, self.width = int(max_batch), int(height), int(width)
        self.num_params = int(self._lib.wg_pc_num_params())
        self.grad_scale = 1.0  # input scaling of the gradient channel (fit_scale)
        self.compress = True  # asinh compression of the gradient channel

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.wg_pc_destroy(h)
            self._h = None

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.num_params, np.float32)
        _lib.check(self._lib.wg_pc_get_params(self._h, ctypes.c_void_p(out.ctypes.data), self.num_params), pc=True)
        return out

    def set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, np.float32)
        if p.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters")
        _lib.check(self._lib.wg_pc_set_params(self._h, ctypes.c_void_p(p.ctypes.data), self.num_params), pc=True)

    def layers(self) -> list[tuple[int, int, int]]:
        """(input channels, output channels, dilation) of each layer."""
        out = []
        for l in range(int(self._lib.wg_pc_num_layers())):
            v = [ctypes.c_int32() for _ in range(3)]
            _lib.check(self._lib.wg_pc_layer_shape(l, *[ctypes.byref(c) for c in v]), pc=True)
            out.append(tuple(c.value for c in v))
        return out

    def fit_scale(self, grads: torch.Tensor) -> float:
        """Sets grad_scale to g / (median per-model RMS of `grads`), with gain
        g = ASINH_GAIN for compressed inputs and 1 otherwise."""
        rms = grads.float().flatten(1).pow(2).mean(1).sqrt()
        gain = ASINH_GAIN if self.compress else 1.0
        self.grad_scale = float(gain / max(float(rms.median()), 1e-30))
        return self.grad_scale

    def save(self, path: str) -> None:
        """Parameters (PyTorch order), then grad_scale and compress, one .npy."""
        extra = np.array([self.grad_scale, 1.0 if self.compress else 0.0], np.float32)
        np.save(path, np.concatenate([self.get_params(), extra]))

    def load(self, path: str) -> None:
        v = np.load(path, allow_pickle=False)
        self.set_params(v[: self.num_params])
        self.grad_scale = float(v[self.num_params])
        self.compress = bool(v[self.num_params + 1] != 0.0)

    def _check_x(self, x: torch.Tensor) -> torch.Tensor:
        if x.dim() != 4 or x.shape[1] != 2 or x.shape[2] != self.height or x.shape[3] != self.width:
            raise ValueError(f"x must be [N, 2, {self.height}, {self.width}]")
        return x.to(device="cuda", dtype=torch.float32).contiguous()

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """[N, 2, H, W] -> [N, H, W] on the device (any N)."""
        x = self._check_x(x)
        y = torch.empty((x.shape[0], self.height, self.width), device="cuda", dtype=torch.float32)
        for s in range(0, x.shape[0], self.max_batch):
            xs, ys = x[s:s + self.max_batch], y[s:s + self.max_batch]
            _lib.check(self._lib.wg_pc_forward_dev(self._h, xs.shape[0], _p(xs), _p(ys), _stream()), pc=True)
        return y

    def loss_grad(self, x: torch.Tensor, target: torch.Tensor) -> tuple[float, np.ndarray]:
        """Mean squared error and its parameter gradient (PyTorch order)."""
        x = self._check_x(x)
        t = target.to(device="cuda", dtype=torch.float32).contiguous()
        loss = np.zeros(1, np.float32)
        grad = np.zeros(self.num_params, np.float32)
        _lib.check(self._lib.wg_pc_loss_grad(self._h, x.shape[0], _p(x), _p(t), ctypes.c_void_p(loss.ctypes.data),
                                             ctypes.c_void_p(grad.ctypes.data), _stream()), pc=True)
        return float(loss[0]), grad

    def train_step(self, x: torch.Tensor, target: torch.Tensor, lr: float, weight_decay: float = 0.0) -> torch.Tensor:
        """One AdamW step; returns the pre-step loss as a device scalar."""
        x = self._check_x(x)
        t = target.to(device="cuda", dtype=torch.float32).contiguous()
        loss = torch.empty(1, device="cuda", dtype=torch.float32)
        _lib.check(self._lib.wg_pc_train_step_dev(self._h, x.shape[0], _p(x), _p(t), float(lr), float(weight_decay),
                                                  _p(loss), _stream()), pc=True)
        return loss

    # -- use in refine ------------------------------------------------------
    def direction(self, grads: torch.Tensor, fields_km: torch.Tensor) -> torch.Tensor:
        """Predicted model update (km/s) for velocity-space gradients."""
        return self.forward(features(grads, fields_km, self.grad_scale, self.compress)) * TARGET_SCALE_KM

    def update_hook(self):
        """refine_batch `update` callable: the predicted model update (km/s)."""

        def update(grads: torch.Tensor, fields_km: torch.Tensor) -> torch.Tensor:
            return self.direction(grads, fields_km)

        return update

    def hook(self):
        """refine_batch `precondition` callable: minus the predicted update,
        rescaled to the norm of the gradient it replaces."""

        def precondition(grads: torch.Tensor, fields_km: torch.Tensor) -> torch.Tensor:
            d = self.direction(grads, fields_km)
            gn = grads.flatten(1).norm(dim=1).clamp_min(1e-30)
            dn = d.flatten(1).norm(dim=1).clamp_min(1e-30)
            return -d * (gn / dn)[:, None, None]

        return precondition


def pseudo_hessian_hook(physics, fields_km: torch.Tensor, eps: float = 0.01):
    """refine_batch `precondition` callable dividing each model's gradient by
    its normalised source illumination P + eps (P = illumination / its max,
    computed once on `fields_km`, the starting fields of the batch): the
    diagonal pseudo-Hessian scaling of the gradient, with no learning.  The
    illumination costs one forward per model."""
    illum = physics.illumination_torch(fields_km)
    P = illum / illum.flatten(1).amax(1).clamp_min(1e-30)[:, None, None]

    def precondition(grads: torch.Tensor, fields: torch.Tensor, ids) -> torch.Tensor:
        return grads / (P[torch.as_tensor(ids, device=grads.device)] + eps)

    return precondition


def make_pairs(physics, n: int, seed: int, bands_hz=(6.0, 9.0, 12.0, 15.0), batch: int = 64,
               net: "Preconditioner | None" = None, offset_km: torch.Tensor | None = None):
    """Training pairs from the batched solver: synthetic (start, true) models
    (the flat / curved / faulted families plus a Gaussian anomaly), the
    velocity-space gradient at the current model in a passband drawn per
    model, and the target (true - current) / 0.3 km/s.  The current model is
    the start model plus `offset_km` ([70, 70], e.g. the field the scene's
    initial Gaussians render; default none), or, with `net`, that model
    moved by the network's predicted update (clipped to 1.5..4.5 km/s):
    pairs at the points a first learned step reaches, so that a second one
    learns to correct the first.
    Returns device tensors grads [n, 70, 70], fields_km [n, 70, 70] and
    target [n, 70, 70]."""
    from . import synthetic

    gs, fs, ts = [], [], []
    rng = np.random.default_rng(seed)
    for s in range(0, n, batch):
        m = min(batch, n - s)
        start, true = synthetic.batch(m, seed=seed * 1000003 + s)
        cur_km = torch.from_numpy(start / 1000.0).cuda()
        if offset_km is not None:
            cur_km = cur_km + offset_km.to(cur_km)
        true_km = torch.from_numpy(true / 1000.0).cuda()
        bands = [bands_hz[k] for k in rng.integers(0, len(bands_hz), m)]
        wav = torch.stack([torch.from_numpy(physics.band_wavelet(b)) for b in bands]).cuda()
        obs = physics.simulate_torch(true_km, wavelets=wav)
        _, grads = physics.loss_and_grad_torch(cur_km, obs, wavelets=wav)
        if net is not None:
            cur_km = (cur_km + net.direction(grads, cur_km)).clamp(1.5, 4.5)
            _, grads = physics.loss_and_grad_torch(cur_km, obs, wavelets=wav)
        gs.append(grads)
        fs.append(cur_km.float())
        ts.append(((true_km - cur_km) / TARGET_SCALE_KM).float())
    return torch.cat(gs), torch.cat(fs), torch.cat(ts)


def train(net: Preconditioner, grads: torch.Tensor, fields_km: torch.Tensor, target: torch.Tensor,
          epochs: int = 20, batch: int = 128, lr: float = 2e-3, weight_decay: float = 1e-4,
          seed: int = 0, fit_scale: bool = True) -> list[float]:
    """Fits grad_scale (unless `fit_scale` is False: a later training stage
    keeps the first stage's input scaling), then mini-batch AdamW with a
    cosine learning-rate decay; returns the mean loss of each epoch."""
    if fit_scale:
        net.fit_scale(grads)
    x = features(grads, fields_km, net.grad_scale, net.compress)
    n = x.shape[0]
    g = torch.Generator(device="cpu").manual_seed(seed)
    steps = epochs * ((n + batch - 1) // batch)
    k = 0
    hist = []
    for _ in range(epochs):
        perm = torch.randperm(n, generator=g).cuda()
        tot = torch.zeros(1, device="cuda")
        nb = 0
        for s in range(0, n, batch):
            idx = perm[s:s + batch]
            rate = lr * 0.5 * (1.0 + np.cos(np.pi * k / max(steps, 1)))
            tot += net.train_step(x[idx], target[idx], rate, weight_decay)
            nb += 1
            k += 1
        hist.append(float(tot) / nb)
    return hist
