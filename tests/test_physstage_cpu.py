"""The phys stage's schedule and balancing (train.py:195-196, 256-268) on the
CPU, with a stand-in physics whose misfit gradient is known in closed form:
the passband schedule against hand-computed tables, and the term against
what its construction fixes: the physics part of the gradient is
phys_ratio * ||g_sup[:nb]|| in norm and parallel to the physics gradient,
rows beyond nb keep the supervised gradient, and the loss grows by
scale * <field, g / nb>."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from wgnet_amd import physstage


def test_parse_bands():
    assert physstage.parse_bands("6,9,12,full") == [6.0, 9.0, 12.0, None]
    assert physstage.parse_bands("5, 8 ,none") == [5.0, 8.0, None]
    assert physstage.parse_bands("15") == [15.0]


def test_band_schedule_tables():
    b = [6.0, 9.0, 12.0, None]
    # int(4 (step - 1) / 10) for steps 1..10 = 0 0 0 1 1 2 2 2 3 3
    assert [physstage.band_at(s, 10, b) for s in range(1, 11)] == [6.0, 6.0, 6.0, 9.0, 9.0, 12.0, 12.0, 12.0,
                                                                    None, None]
    # one step per band, and past the end the last band holds
    assert [physstage.band_at(s, 4, b) for s in range(1, 7)] == [6.0, 9.0, 12.0, None, None, None]
    assert physstage.band_at(1, 0, b) == 6.0 and physstage.band_at(3, 0, b) is None


class _StandIn:
    """misfit_b(f) = 0.5 ||A f_b - d_b||^2 with A = diag(a): its gradient is
    a * (a f_b - d_b); observed d_b = a * true_b."""

    def __init__(self, n=70):
        rng = np.random.default_rng(5)
        self.a = torch.from_numpy((0.5 + rng.random((n, n))).astype(np.float32))
        self.fmax, self.bands_set, self.sim_in, self.lg_in = None, [], [], []

    def set_band(self, f):
        self.fmax = f
        self.bands_set.append(f)

    def simulate_torch(self, true_km):
        self.sim_in.append(true_km.clone())
        return self.a * true_km

    def loss_and_grad_true_torch(self, field_km, true_km):
        assert not field_km.requires_grad
        observed = self.simulate_torch(true_km)
        self.lg_in.append(field_km.clone())
        r = self.a * field_km - observed
        return 0.5 * (r * r).sum((1, 2)), self.a * r, observed


def _setup(B=5, seed=0):
    rng = np.random.default_rng(seed)
    w = torch.from_numpy(rng.standard_normal((B, 70, 70)).astype(np.float32)).requires_grad_(True)
    true = torch.from_numpy((2.0 + rng.random((B, 70, 70))).astype(np.float32))
    field = 2.5 + 0.5 * torch.tanh(w)
    loss = (field - true).abs().mean()
    return w, true, field, loss


@pytest.mark.parametrize("nb", [1, 3, 5, 9])
def test_physics_term_balancing(nb):
    phys = _StandIn()
    w, true, field, loss = _setup()
    B = field.shape[0]
    k = min(nb, B)
    (g_sup,) = torch.autograd.grad(loss, field, retain_graph=True)
    new, parts = physstage.physics_term(phys, loss, field, true, nb, phys_ratio=0.5)
    assert torch.equal(phys.sim_in[0], true[:k]) and torch.equal(phys.lg_in[0], field[:k].detach())
    (g_new,) = torch.autograd.grad(new, field)
    g = (phys.a * (phys.a * field[:k].detach() - phys.a * true[:k])).double()
    d = g_new.double()[:k] - g_sup.double()[:k]  # the physics part of the gradient
    assert torch.equal(g_new[k:], g_sup[k:])
    # norm fixed by the balancing, direction by the physics gradient; 1e-4 is
    # the rounding of the float32 norms over up to 24,500 terms
    assert abs(float(d.norm()) - 0.5 * float(g_sup[:k].double().norm())) <= 1e-4 * float(d.norm())
    cos = float((d * g).sum() / (d.norm() * g.norm()))
    assert cos > 1.0 - 1e-6
    s = 0.5 * float(g_sup[:k].double().norm()) / float((g / k).norm())
    assert abs(float(parts["phys_scale"]) - s) <= 1e-4 * s
    added = s * float((field[:k].detach().double() * g / k).sum())
    assert abs(float(new) - float(loss) - added) <= 1e-4 * abs(added) + 1e-6
    assert abs(float(parts["phys_misfit"]) - float(0.5 * ((phys.a * (field[:k].detach() - true[:k])) ** 2)
                                                    .double().sum((1, 2)).mean())) <= 1e-5 * float(parts["phys_misfit"])


def test_phys_stage_sets_each_band_once():
    phys = _StandIn()
    stage = physstage.PhysStage(steps=8, bands="6,9,12,full", phys_ratio=0.5, phys_batch=2, physics=phys)
    seen = []
    for step in range(1, 9):
        w, true, field, loss = _setup(B=3, seed=step)
        new, parts = stage(loss, field, true, step)
        new.backward()
        seen.append(parts["band_hz"])
        assert w.grad is not None and bool(torch.isfinite(w.grad).all())
    assert seen == [6.0, 6.0, 9.0, 9.0, 12.0, 12.0, 15.0, 15.0]
    assert phys.bands_set == [6.0, 9.0, 12.0, None]


def test_physics_term_rejects_empty_batch():
    phys = _StandIn()
    w, true, field, loss = _setup()
    with pytest.raises(ValueError):
        physstage.physics_term(phys, loss, field, true, 0)
