"""GPU parity: the HIP engine against the CPU oracle on the same inputs.

Tolerances (written per test):
  gathers          rel-L2 <= 1e-5 vs the float32 oracle (the kernels keep the
                   reference's operation order, so the expected error is 0)
  adjoint / loss   rel-L2 <= 1e-6 vs the float32 oracle's evaluate_adjoint
  dot products     |<Fx,y> - <x,F^T y>| <= 1e-6 |<Fx,y>| with y = F x, and
                   <= 1e-5 with y random (float32 roundoff of F, the same
                   level as the float32 CPU oracle); FP64 accumulation
  finite diffs     reference test (wave_wgpu.rs:1056-1108) < 3 %; per-cell
                   gradient vs FP64 central differences < 1e-3
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SURFACE_SRC = np.array([[0, 10], [170, 10], [340, 10], [520, 10], [690, 10]], np.float32)


def relerr(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def three_layers(seed: int) -> np.ndarray:
    """3 flat layers: interfaces at U{10..60} cells, velocities U[1500, 4500]
    sorted increasing with depth."""
    rng = np.random.default_rng(seed)
    d = np.sort(rng.integers(10, 61, 2))
    v = np.sort(rng.uniform(1500, 4500, 3)).astype(np.float32)
    m = np.full((70, 70), v[0], np.float32)
    m[d[0]:] = v[1]
    m[d[1]:] = v[2]
    return m


def gaussian_bumps():
    """Inputs of the reference's adjoint FD test (wave_wgpu.rs:1058-1074)."""
    x, z = np.meshgrid(np.arange(70, dtype=np.float32), np.arange(70, dtype=np.float32))
    g1 = np.exp(-((x - 8) ** 2 + (z - 9) ** 2) / 50.0)
    g2 = np.exp(-((x - 56) ** 2 + (z - 10) ** 2) / 50.0)
    true = (2500.0 + 320.0 * g1 - 230.0 * g2).astype(np.float32)
    cur = (2500.0 + 180.0 * g1 - 120.0 * g2).astype(np.float32)
    dirn = (np.exp(-((x - 7) ** 2 + (z - 7) ** 2) / 150.0) + 0.3 * np.exp(-((x - 58) ** 2 + (z - 12) ** 2) / 100.0)).astype(np.float32)
    return true, cur, dirn


@pytest.fixture(scope="module")
def W():
    from wgnet_amd import wave_gauss_rs

    return wave_gauss_rs


def oracle_gathers(orc, vel, src, nt, dt, wavelet, f64=False):
    out = []
    for s in src:
        g = orc.Geometry(nt=nt, dt=dt, src_x=float(s[0]), src_z=float(s[1]))
        out.append(orc.forward_f64(g, vel, wavelet.astype(np.float64)) if f64 else orc.forward_f32(g, vel, wavelet))
    return np.stack(out)


def test_forward_70x70x1000_5shots_matches_cpu_solver(W, orc):
    """3 flat layers, 5 shots at x-cells {0,17,34,52,69}, 1 ms x 1000."""
    E = W.PyWaveEngine(5, 1000, 0.001, 15.0, 20, 1000.0, 5000.0)
    E.set_sources(SURFACE_SRC)
    vel = three_layers(0)
    gpu = E.simulate(vel.reshape(-1))
    assert gpu.shape == (5, 1000, 70) and gpu[:, ::4].shape == (5, 250, 70)
    ref32 = oracle_gathers(orc, vel, SURFACE_SRC, 1000, 0.001, E.wavelet())
    ref64 = oracle_gathers(orc, vel, SURFACE_SRC, 1000, 0.001, E.wavelet(), f64=True)
    e32, e64 = relerr(gpu, ref32), relerr(gpu, ref64)
    print(f"70x70x1000 gathers: rel-L2 vs FP32 CPU {e32:.3e} (bit-equal {np.mean(gpu == ref32):.4f}), vs FP64 CPU {e64:.3e}")
    assert e32 <= 1e-5
    assert e64 <= 1e-4


def test_reference_two_layer_kat(W, orc):
    """wave_wgpu.rs:1023-1054: 2500 over 3200 m/s below z=35, nt=2001 at 0.5 ms,
    shot 0 vs the CPU solver; the reference accepts 5e-4."""
    E = W.PyWaveEngine(5, 2001, 0.0005, 15.0, 20, 1000.0, 5000.0)
    vel = np.full((70, 70), 2500.0, np.float32)
    vel[35:] = 3200.0
    gpu = E.simulate(vel.reshape(-1))
    cpu = W.simulate_shot_cpu  # the drop-in single-shot solver (GPU-backed)
    ref = oracle_gathers(orc, vel, [[0.0, 10.0]], 2001, 0.0005, E.wavelet())[0]
    assert relerr(gpu[0], ref) < 5e-4
    assert relerr(gpu[0], ref) <= 1e-5
    one = cpu(vel.reshape(-1), 10.0, 0.0005, 2001, 20, 1000.0, 5000.0, 15.0, 0.0, 10.0)
    assert one.shape == (2001, 70)
    assert relerr(one, ref) <= 1e-5


def test_other_grid_runtime_kernels(W, orc):
    """A grid other than 70 x 70 / pml 20 runs the runtime-geometry kernels:
    gathers and gradient must equal the oracle's on it too."""
    nvx, nvz, pml, nt = 64, 48, 12, 300
    E = W.PyWaveEngine(2, nt, 0.001, 15.0, pml, 1000.0, 5000.0, nvx=nvx, nvz=nvz)
    E.set_sources(np.array([[31.0, 10.0], [407.5, 24.0]], np.float32))
    rng = np.random.default_rng(11)
    true = np.full((nvz, nvx), 2000.0, np.float32)
    true[12:] = 2800.0
    true[30:] = 3600.0
    cur = true + rng.uniform(-80, 80, true.shape).astype(np.float32)
    obs = E.simulate(true.reshape(-1))
    for s, (x, z) in enumerate(E.sources()):
        g = orc.Geometry(nvx=nvx, nvz=nvz, pml=pml, nt=nt, dt=0.001, src_x=float(x), src_z=float(z))
        assert relerr(obs[s], orc.forward_f32(g, true, E.wavelet())) <= 1e-5
    loss, grad = E.evaluate_adjoint(cur.reshape(-1), obs.reshape(-1))
    g = orc.Geometry(nvx=nvx, nvz=nvz, pml=pml, nt=nt, dt=0.001)
    rl, rg = orc.evaluate_adjoint_f32(g, E.sources(), cur, E.wavelet(), obs)
    print(f"64x48 pml12: loss {loss:.9e} vs {rl:.9e}, grad rel-L2 {relerr(grad, rg):.3e}")
    assert abs(loss - rl) <= 1e-6 * rl
    assert relerr(grad, rg) <= 1e-6


def test_default_sources_and_setters(W):
    E = W.PyWaveEngine(5, 50)
    np.testing.assert_allclose(E.sources(), [[0, 10], [172.5, 10], [345, 10], [517.5, 10], [690, 10]])
    assert (E.shots, E.nt, E.nrec) == (5, 50, 70)
    E1 = W.PyWaveEngine(1, 50)
    np.testing.assert_allclose(E1.sources(), [[345, 10]])
    with pytest.raises(ValueError, match="shape"):
        E.set_sources(np.zeros((4, 2), np.float32))
    with pytest.raises(ValueError, match="nt"):
        E.set_wavelet(np.zeros(49, np.float32))
    with pytest.raises(ValueError, match="4900"):
        E.simulate(np.zeros(4899, np.float32))
    with pytest.raises(ValueError, match="observed"):
        E.evaluate_adjoint(np.full(4900, 2000, np.float32), np.zeros(10, np.float32))
    with pytest.raises(TypeError):
        E.simulate(np.zeros(4900, np.float64))


def test_evaluate_adjoint_matches_oracle(W, orc):
    nt = 400
    E = W.PyWaveEngine(3, nt, 0.001, 15.0, 20, 1000.0, 5000.0)
    E.set_sources(np.array([[55.0, 10.0], [333.3, 17.5], [612.0, 10.0]], np.float32))  # off-node sources
    true = three_layers(1)
    cur = true.copy()
    cur[30:40, 20:45] += 120.0
    cur[0, 0], cur[69, 69] = 900.0, 5200.0  # clamped cells
    obs = E.simulate(true.reshape(-1))
    loss, grad = E.evaluate_adjoint(cur.reshape(-1), obs.reshape(-1))
    g = orc.Geometry(nt=nt, dt=0.001)
    rl, rg = orc.evaluate_adjoint_f32(g, E.sources(), cur, E.wavelet(), obs)
    print(f"loss gpu {loss:.9e} oracle {rl:.9e}; grad rel-L2 {relerr(grad, rg):.3e}")
    assert abs(loss - rl) <= 1e-6 * rl
    assert relerr(grad, rg) <= 1e-6
    assert grad[0, 0] == 0.0 and grad[69, 69] == 0.0  # clamp mask
    l64, g64 = orc.evaluate_adjoint_f64(g, E.sources(), cur, E.wavelet().astype(np.float64), obs.astype(np.float64))
    print(f"vs FP64: loss rel {abs(loss - l64) / l64:.3e}, grad rel-L2 {relerr(grad, g64):.3e}")
    assert relerr(grad, g64) <= 1e-3


def test_zero_energy_raises(W):
    E = W.PyWaveEngine(2, 40, 0.001)
    with pytest.raises(ValueError, match="zero energy"):
        E.evaluate_adjoint(np.full(4900, 2000, np.float32), np.zeros(2 * 40 * 70, np.float32))


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _dot(a, b):
    return float((a.double() * b.double()).sum())


def test_dot_product_wavelet_operator(W):
    """F: per-shot wavelets -> gathers at fixed v; F^T y from the adjoint.
    Gate 1e-6 with y = F x (no cancellation in <Fx, y>); with y random the
    float32 roundoff of F sets the level (~1e-6), gated at 1e-5."""
    import ctypes

    from wgnet_amd import _lib
    from wgnet_amd.physics import _stream

    B, S, nt = 2, 3, 500
    E = W.PyWaveEngine(S, nt, 0.001, 15.0, 20, 1000.0, 5000.0)
    rng = np.random.default_rng(7)
    vel = _dev(np.stack([three_layers(2), three_layers(3)]))
    P = lambda t: ctypes.c_void_p(t.data_ptr())

    def F(x):
        out = torch.empty((B, S, nt, 70), device="cuda")
        _lib.check(E._lib.wg_simulate_wavelets_dev(E.handle, B, P(vel), P(x), P(out), _stream()))
        return out

    def Ft(y):
        out = torch.empty((B, S, nt), device="cuda")
        _lib.check(E._lib.wg_adjoint_apply_dev(E.handle, B, P(vel), P(y), None, P(out), None, _stream()))
        return out

    x = _dev(rng.standard_normal((B, S, nt)).astype(np.float32))
    for name, y, tol in [("y = F x", F(x), 1e-6),
                         ("y random", _dev(rng.standard_normal((B, S, nt, 70)).astype(np.float32)), 1e-5)]:
        lhs, rhs = _dot(F(x), y), _dot(x, Ft(y.contiguous()))
        rel = abs(lhs - rhs) / abs(lhs)
        print(f"wavelet dot test ({name}): <Fx,y>={lhs:.12e} <x,F^T y>={rhs:.12e} rel={rel:.3e}")
        assert rel <= tol


def test_dot_product_born_operator(W, orc):
    """J = d(gathers)/d(v): Born forward vs the adjoint gradient.  Gate 1e-6 on
    <J dv, J dv> = <dv, J^T J dv> (well conditioned); random y is limited by
    float32 roundoff of J (the CPU float32 oracle shows 0.4-3e-6 on the same
    test), gated at 1e-5."""
    import ctypes

    from wgnet_amd import _lib
    from wgnet_amd.physics import _stream

    B, S, nt = 2, 2, 500
    E = W.PyWaveEngine(S, nt, 0.001, 15.0, 20, 1000.0, 5000.0)
    E.set_sources(np.array([[123.0, 10.0], [456.7, 33.3]], np.float32))
    rng = np.random.default_rng(8)
    v = np.stack([three_layers(4), three_layers(5)])
    v[0, 3, 3], v[1, 50, 60] = 800.0, 5400.0
    vel = _dev(v)
    P = lambda t: ctypes.c_void_p(t.data_ptr())

    def J(dv):
        out = torch.empty((B, S, nt, 70), device="cuda")
        _lib.check(E._lib.wg_born_dev(E.handle, B, P(vel), P(dv), P(out), _stream()))
        return out

    def Jt(y):
        out = torch.empty((B, 70, 70), device="cuda")
        _lib.check(E._lib.wg_adjoint_apply_dev(E.handle, B, P(vel), P(y), P(out), None, None, _stream()))
        return out

    dv = _dev(rng.standard_normal((B, 70, 70)).astype(np.float32))
    jdv = J(dv)
    for name, y, tol in [("y = J dv", jdv, 1e-6),
                         ("y random", _dev(rng.standard_normal((B, S, nt, 70)).astype(np.float32)), 1e-5)]:
        lhs, rhs = _dot(jdv, y), _dot(dv, Jt(y.contiguous()))
        rel = abs(lhs - rhs) / abs(lhs)
        print(f"Born dot test ({name}): <J dv,y>={lhs:.12e} <dv,J^T y>={rhs:.12e} rel={rel:.3e}")
        assert rel <= tol
    # the GPU Born gathers against the FP64 oracle's tangent-linear model
    g = orc.Geometry(nt=nt, dt=0.001, src_x=123.0, src_z=10.0)
    ref = orc.born_f64(g, v[0], dv[0].cpu().numpy().astype(np.float64), E.wavelet().astype(np.float64))
    err = relerr(jdv[0, 0].cpu().numpy(), ref)
    print(f"Born gathers vs FP64 oracle: rel-L2 {err:.3e}")
    assert err <= 1e-4


def test_reference_adjoint_fd(W):
    """wave_wgpu.rs:1056-1108: 2 shots, nt=501, eps = 1 m/s along `dir`, < 3 %."""
    E = W.PyWaveEngine(2, 501, 0.0005, 15.0, 20, 1000.0, 5000.0)
    true, cur, dirn = gaussian_bumps()
    obs = E.simulate(true.reshape(-1)).reshape(-1)
    _, grad = E.evaluate_adjoint(cur.reshape(-1), obs)
    lp, _ = E.evaluate_adjoint((cur + dirn).reshape(-1), obs)
    lm, _ = E.evaluate_adjoint((cur - dirn).reshape(-1), obs)
    fd = (lp - lm) / 2.0
    adj = float((grad.astype(np.float64) * dirn).sum())
    print(f"reference FD test: fd={fd:.6e} adj={adj:.6e} rel={abs(fd - adj) / abs(fd):.3e}")
    assert abs(fd - adj) / abs(fd) < 0.03


def test_per_cell_gradient_vs_fp64_fd(W, orc):
    """Central differences at h = 1 m/s on single cells, evaluated with the FP64
    oracle (FP32 losses cannot resolve them), against the GPU gradient."""
    nt = 400
    E = W.PyWaveEngine(2, nt, 0.001, 15.0, 20, 1000.0, 5000.0)
    true = three_layers(6)
    yy, xx = np.mgrid[0:70, 0:70]
    true = (true + 250.0 * np.exp(-((xx - 35) ** 2 + (yy - 30) ** 2) / (2 * 5.0 ** 2))).astype(np.float32)
    cur = three_layers(6)
    obs = E.simulate(true.reshape(-1))
    _, grad = E.evaluate_adjoint(cur.reshape(-1), obs.reshape(-1))
    g = orc.Geometry(nt=nt, dt=0.001)
    src = E.sources()
    w64 = E.wavelet().astype(np.float64)
    obs64 = np.stack([orc.forward_f64(orc.Geometry(nt=nt, dt=0.001, src_x=s[0], src_z=s[1]), true, w64) for s in src])
    rng = np.random.default_rng(9)
    cells = [(30, 35), (20, 10)] + [tuple(rng.integers(5, 65, 2)) for _ in range(2)]
    for (z, x) in cells:
        vp, vm = cur.copy(), cur.copy()
        vp[z, x] += 1.0
        vm[z, x] -= 1.0
        lp, _ = orc.evaluate_adjoint_f64(g, src, vp, w64, obs64)
        lm, _ = orc.evaluate_adjoint_f64(g, src, vm, w64, obs64)
        fd = (lp - lm) / 2.0
        print(f"cell {(z, x)}: fd={fd:.6e} gpu={grad[z, x]:.6e}")
        assert abs(fd - grad[z, x]) <= 1e-3 * abs(fd) + 1e-3 * np.abs(grad).max()


def test_batched_device_path_equals_host_path(W):
    from wgnet_amd.physics import WavePhysics

    ph = WavePhysics(shots=3, nt=300, dt=0.001)
    fields = np.stack([three_layers(s) / 1000.0 for s in (10, 11, 12)]).astype(np.float32)
    obs = ph.simulate(fields + 0.05)
    losses, grads = ph.loss_and_grad(fields, obs)
    for b in range(3):
        l, gr = ph.engine.evaluate_adjoint(np.ascontiguousarray(fields[b].reshape(-1) * 1000.0, dtype=np.float32),
                                           np.ascontiguousarray(obs[b].reshape(-1)))
        assert abs(losses[b] - l) <= 1e-6 * l
        assert relerr(grads[b], gr * 1000.0) <= 1e-6
    # autograd: backward of sum(losses * c) is c_b * grad_b
    f = torch.from_numpy(fields).cuda().requires_grad_(True)
    c = torch.tensor([1.0, -2.0, 0.5], device="cuda")
    (ph.misfit(f, torch.from_numpy(obs).cuda()) * c).sum().backward()
    ref = torch.from_numpy(grads).cuda() * c.view(-1, 1, 1)
    assert torch.allclose(f.grad, ref, rtol=1e-6, atol=0)


def test_band_identity(W):
    """verify.py:76-84: simulate with the low-passed wavelet == low-pass of the
    full-band data (reference tolerance 2e-3 max-rel)."""
    from wgnet_amd.physics import WavePhysics

    ph = WavePhysics(shots=2, nt=501, dt=0.0005)
    fields = (three_layers(13) / 1000.0)[None]
    ph.set_band(8.0)
    low = ph.simulate(fields)
    ph.set_band(None)
    full = ph.simulate(fields)
    ph.set_band(8.0)
    filt = ph.bandlimit(full)
    ph.set_band(None)
    rel = float(np.abs(low - filt).max() / np.abs(filt).max())
    print(f"band identity max-rel {rel:.3e}")
    assert rel < 2e-3


@pytest.mark.parametrize("B", [64, 130])
def test_overlapped_chunk_pipeline_is_bit_identical(B):
    """Chunks of one unit per CU on two streams (the default) give the same
    losses and gradients as the chunks run one after another."""
    import ctypes

    import torch

    from wgnet_amd import synthetic
    from wgnet_amd.physics import WavePhysics

    ph = WavePhysics(shots=5, nt=300, dt=0.001)
    ph.engine._lib.wg_set_time_split(ph.engine.handle, 0)  # the chunk paths, not the work queue
    start, true = synthetic.batch(B, seed=4242)
    obs = ph.simulate_torch(torch.from_numpy(true / 1000.0).cuda())
    f = torch.from_numpy(start / 1000.0).cuda()
    l_on, g_on = ph.loss_and_grad_torch(f, obs)
    ph.engine._lib.wg_set_overlap(ph.engine.handle, 0)
    l_off, g_off = ph.loss_and_grad_torch(f, obs)
    ph.engine._lib.wg_set_overlap(ph.engine.handle, 1)
    torch.cuda.synchronize()
    assert torch.equal(l_on, l_off)
    assert torch.equal(g_on, g_off)


@pytest.mark.parametrize("B,shots,nt", [(52, 5, 301), (64, 5, 300), (130, 2, 257), (257, 1, 257)])
def test_time_split_work_queue_is_bit_identical(B, shots, nt):
    """The resident work queue (each unit's forward and adjoint cut into 1,
    2, 3, 4 or 7 time parts, state handed over through HBM, one launch for
    the batch) gives the same traces, losses and gradients as whole-unit
    chunk launches.  Every shape has more units than the device has compute
    units (the queues step aside otherwise), with an uneven last round."""
    import ctypes

    import torch

    from wgnet_amd import synthetic
    from wgnet_amd.physics import WavePhysics

    ph = WavePhysics(shots=shots, nt=nt, dt=0.001)
    lib, h = ph.engine._lib, ph.engine.handle
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    assert B * shots > cus, (B * shots, cus)
    start, true = synthetic.batch(B, seed=4343 + B)
    obs = ph.simulate_torch(torch.from_numpy(true / 1000.0).cuda())
    f = torch.from_numpy(start / 1000.0).cuda()
    R = ph.engine.nrec
    lib.wg_set_time_split(h, 0)
    tr0 = torch.empty((B, shots, nt, R), device="cuda")
    l0, g0 = ph.loss_and_grad_torch(f, obs, traces_out=tr0)
    assert bool(torch.isfinite(l0).all()) and float(g0.abs().max()) > 0.0
    fms, fn, ams, an = ctypes.c_double(), ctypes.c_int64(), ctypes.c_double(), ctypes.c_int64()
    for parts in (1, 2, 3, 4, 7):
        lib.wg_set_time_split(h, parts)
        tr = torch.empty_like(tr0)
        lib.wg_timing_enable(h, 1)
        lib.wg_timing_read(h, ctypes.byref(fms), ctypes.byref(fn), ctypes.byref(ams), ctypes.byref(an))
        l, g = ph.loss_and_grad_torch(f, obs, traces_out=tr)
        torch.cuda.synchronize()
        lib.wg_timing_read(h, ctypes.byref(fms), ctypes.byref(fn), ctypes.byref(ams), ctypes.byref(an))
        lib.wg_timing_enable(h, 0)
        assert (fn.value, an.value) == (1, 1), (parts, fn.value, an.value)  # the queues ran
        assert torch.equal(tr, tr0), parts
        assert torch.equal(l, l0), parts
        assert torch.equal(g, g0), parts


def test_illumination_matches_oracle_snapshots(orc):
    """wg_illumination_dev: sum over shots and steps of the squared snapshot at
    each visible cell, against the oracle's snapshots summed in float64."""
    import torch

    from wgnet_amd.physics import WavePhysics

    nt = 300
    ph = WavePhysics(shots=2, nt=nt, dt=0.001)
    src = np.array([[100.0, 10.0], [555.0, 10.0]], np.float32)
    ph.engine.set_sources(src)
    vel = np.full((70, 70), 1800.0, np.float32)
    vel[30:] = 2600.0
    vel[50:] = 3400.0
    il = ph.illumination_torch(torch.from_numpy(vel[None] / 1000.0).cuda())[0].cpu().numpy()
    ref = np.zeros((70, 70))
    for s in range(2):
        g = orc.Geometry(nt=nt, dt=0.001, src_x=float(src[s, 0]), src_z=float(src[s, 1]))
        _, snaps = orc.forward_f32(g, vel, ph.engine.wavelet(), snaps=True)
        d = g.dims()
        full = np.asarray(snaps, np.float64).reshape(nt, d.nz, d.nx)  # [nt, padded cells]
        ref += (full[:, d.pad:d.pad + 70, d.pad:d.pad + 70] ** 2).sum(0)
    err = np.linalg.norm(il - ref) / np.linalg.norm(ref)
    print(f"illumination rel-L2 vs oracle {err:.2e}")
    assert err <= 1e-6
