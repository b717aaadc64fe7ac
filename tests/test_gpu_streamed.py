"""GPU parity of the HBM-streamed kernels (grids too large for one workgroup)
against the CPU oracle, with and without checkpoint recompute."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def layered(nvz, nvx, seed):
    rng = np.random.default_rng(seed)
    m = np.full((nvz, nvx), 1800.0, np.float32)
    for k, z in enumerate(np.sort(rng.integers(nvz // 6, nvz - 4, 3))):
        m[z:] = 2200.0 + 600.0 * k
    return m


@pytest.mark.parametrize("fwd_block", [1, 2])
def test_streamed_100x100_matches_oracle(orc, fwd_block):
    """Config 2's 100 x 100 grid (144 x 144 padded) on the HBM-streamed
    kernels (the row bands it runs on by default are tested in
    test_gpu_bands.py), with one and two forward steps per launch."""
    from wgnet_amd.wave_gauss_rs import PyWaveEngine

    nt = 300
    E = PyWaveEngine(2, nt, 0.001, 15.0, 20, 1000.0, 5000.0, nvx=100, nvz=100)
    assert E.streamed and E.bands == 2
    E.set_band_split(False)
    E.set_forward_time_block(fwd_block)
    E.set_sources(np.array([[150.0, 10.0], [777.7, 10.0]], np.float32))
    true = layered(100, 100, 1)
    cur = true.copy()
    cur[20:40, 30:70] += 150.0
    obs = E.simulate(true.reshape(-1))
    for s, (x, z) in enumerate(E.sources()):
        g = orc.Geometry(nvx=100, nvz=100, nt=nt, dt=0.001, src_x=float(x), src_z=float(z))
        err = relerr(obs[s], orc.forward_f32(g, true, E.wavelet()))
        print(f"100x100 shot {s}: gather rel-L2 {err:.3e}")
        assert err <= 1e-5
    loss, grad = E.evaluate_adjoint(cur.reshape(-1), obs.reshape(-1))
    rl, rg = orc.evaluate_adjoint_f32(orc.Geometry(nvx=100, nvz=100, nt=nt, dt=0.001), E.sources(), cur,
                                      E.wavelet(), obs)
    print(f"100x100: loss {loss:.9e} vs {rl:.9e}; grad rel-L2 {relerr(grad, rg):.3e}")
    assert abs(loss - rl) <= 1e-6 * rl
    assert relerr(grad, rg) <= 1e-6


def test_streamed_checkpoint_recompute_is_exact(orc):
    """A 300 x 100 grid (config 3's shape, reduced) at dt = 0.8 ms: the
    checkpoint-every-100 adjoint equals the stored-snapshot adjoint bit for
    bit, and both equal the oracle."""
    from wgnet_amd.wave_gauss_rs import PyWaveEngine

    nt, nvx, nvz = 450, 300, 100
    E = PyWaveEngine(3, nt, 0.0008, 15.0, 20, 1000.0, 5000.0, nvx=nvx, nvz=nvz)
    assert E.streamed and E.bands == 0
    E.set_sources(np.array([[0.0, 10.0], [1495.0, 10.0], [2990.0, 10.0]], np.float32))
    true = layered(nvz, nvx, 2)
    yy, xx = np.mgrid[0:nvz, 0:nvx]
    true = np.clip(true + 300.0 * np.exp(-((xx - 150) ** 2 + (yy - 30) ** 2) / (2 * 12.0 ** 2)), 1500, 4500).astype(np.float32)
    cur = layered(nvz, nvx, 2)
    obs = E.simulate(true.reshape(-1))
    loss_a, grad_a = E.evaluate_adjoint(cur.reshape(-1), obs.reshape(-1))
    E.set_snapshot_budget(64 << 20)  # forces checkpoints every 100 steps
    loss_b, grad_b = E.evaluate_adjoint(cur.reshape(-1), obs.reshape(-1))
    assert loss_a == loss_b
    assert np.array_equal(grad_a, grad_b)
    rl, rg = orc.evaluate_adjoint_f32(orc.Geometry(nvx=nvx, nvz=nvz, nt=nt, dt=0.0008), E.sources(), cur,
                                      E.wavelet(), obs)
    print(f"300x100: loss {loss_b:.9e} vs {rl:.9e}; grad rel-L2 {relerr(grad_b, rg):.3e}")
    assert abs(loss_b - rl) <= 1e-6 * rl
    assert relerr(grad_b, rg) <= 1e-6


@pytest.mark.parametrize("bands", [False, True])
def test_streamed_dot_products(bands):
    """Wavelet- and Born-operator dot tests on the streamed kernels (y = F x);
    with the row bands on, the wavelet forward runs on them."""
    import ctypes

    import torch

    from wgnet_amd import _lib
    from wgnet_amd.physics import _stream
    from wgnet_amd.wave_gauss_rs import PyWaveEngine

    B, S, nt = 1, 2, 300
    E = PyWaveEngine(S, nt, 0.001, 15.0, 20, 1000.0, 5000.0, nvx=100, nvz=100)
    E.set_band_split(bands)
    rng = np.random.default_rng(3)
    P = lambda t: ctypes.c_void_p(t.data_ptr())
    vel = torch.from_numpy(layered(100, 100, 5)[None].copy()).cuda()
    x = torch.from_numpy(rng.standard_normal((B, S, nt)).astype(np.float32)).cuda()
    fx = torch.empty((B, S, nt, 100), device="cuda")
    _lib.check(E._lib.wg_simulate_wavelets_dev(E.handle, B, P(vel), P(x), P(fx), _stream()))
    ft = torch.empty((B, S, nt), device="cuda")
    _lib.check(E._lib.wg_adjoint_apply_dev(E.handle, B, P(vel), P(fx), None, P(ft), None, _stream()))
    lhs, rhs = float((fx.double() ** 2).sum()), float((x.double() * ft.double()).sum())
    print(f"streamed wavelet dot: rel {abs(lhs - rhs) / abs(lhs):.3e}")
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)
    dv = torch.from_numpy(rng.standard_normal((B, 100, 100)).astype(np.float32)).cuda()
    jdv = torch.empty((B, S, nt, 100), device="cuda")
    _lib.check(E._lib.wg_born_dev(E.handle, B, P(vel), P(dv), P(jdv), _stream()))
    gr = torch.empty((B, 100, 100), device="cuda")
    _lib.check(E._lib.wg_adjoint_apply_dev(E.handle, B, P(vel), P(jdv), P(gr), None, None, _stream()))
    lhs, rhs = float((jdv.double() ** 2).sum()), float((dv.double() * gr.double()).sum())
    print(f"streamed Born dot: rel {abs(lhs - rhs) / abs(lhs):.3e}")
    assert abs(lhs - rhs) <= 1e-6 * abs(lhs)


@pytest.mark.parametrize("nvx,nvz,nt,budget", [(100, 100, 301, None), (150, 90, 251, 40 << 20)])
def test_time_blocked_adjoint_is_bit_identical(nvx, nvz, nt, budget):
    """Two adjoint steps per launch (tile + recomputed 2-cell ring) equal
    single steps bit for bit, on every tile shape: an odd step count (one
    single step at the end), a grid that is not a multiple of any tile, and
    checkpoint segments of 100, 100 and 51 steps; receivers, sources and
    the C-PML strips cross tile edges."""
    from wgnet_amd.wave_gauss_rs import PyWaveEngine

    E = PyWaveEngine(3, nt, 0.001, 15.0, 20, 1000.0, 5000.0, nvx=nvx, nvz=nvz)
    assert E.streamed
    E.set_band_split(False)
    if budget:
        E.set_snapshot_budget(budget)
    w = (nvx - 1) * 10.0
    E.set_sources(np.array([[0.0, 10.0], [0.37 * w, 35.0], [w, 150.0]], np.float32))
    true = layered(nvz, nvx, 5)
    cur = true.copy()
    cur[nvz // 3:nvz // 2, nvx // 4:nvx // 2] += 200.0
    obs = E.simulate(true.reshape(-1)).reshape(-1)
    results = {}
    for tile in range(4):
        E.set_stream_tile(tile)
        for tb in (1, 2):
            E.set_time_block(tb)
            results[(tile, tb)] = E.evaluate_adjoint(cur.reshape(-1), obs)
    ref_loss, ref_grad = results[(0, 1)]
    assert np.abs(ref_grad).max() > 0.0
    for key, (loss, grad) in results.items():
        assert loss == ref_loss, key
        assert np.array_equal(grad, ref_grad), key


@pytest.mark.parametrize("nvx,nvz,nt,budget", [(100, 100, 301, None), (100, 100, 300, None),
                                               (150, 90, 251, 40 << 20), (150, 90, 250, 40 << 20)])
def test_time_blocked_forward_is_bit_identical(nvx, nvz, nt, budget):
    """Two forward steps per launch (tile + recomputed 2-cell ring, the
    second step's gathers recorded by the next launch) equal single steps bit
    for bit, on every tile shape and with both adjoint blockings: gathers
    (odd and even step counts, so the last gathers come from a two-step
    launch's successor or from a record launch before a final single step),
    and loss and gradient, through stored snapshots or checkpoint segments
    (100, 100, 51 / 50 steps: a segment of odd length ends with a single
    step); receivers, sources and the C-PML strips cross tile edges."""
    from wgnet_amd.wave_gauss_rs import PyWaveEngine

    E = PyWaveEngine(3, nt, 0.001, 15.0, 20, 1000.0, 5000.0, nvx=nvx, nvz=nvz)
    assert E.streamed
    E.set_band_split(False)
    E.set_forward_time_block(1)  # the reference: single forward steps
    if budget:
        E.set_snapshot_budget(budget)
    w = (nvx - 1) * 10.0
    E.set_sources(np.array([[0.0, 10.0], [0.37 * w, 35.0], [w, 150.0]], np.float32))
    true = layered(nvz, nvx, 6)
    cur = true.copy()
    cur[nvz // 4:nvz // 2, nvx // 3:nvx // 2] -= 150.0
    obs_ref = E.simulate(true.reshape(-1))
    assert np.abs(obs_ref[:, -1]).max() > 0.0  # the last step's gathers are not trivially zero
    ref = E.evaluate_adjoint(cur.reshape(-1), obs_ref.reshape(-1))
    assert np.abs(ref[1]).max() > 0.0
    E.set_forward_time_block(2)
    for tile in range(4):
        E.set_stream_tile(tile)
        obs = E.simulate(true.reshape(-1))
        assert np.array_equal(obs, obs_ref), tile
        for tb in (1, 2):
            E.set_time_block(tb)
            loss, grad = E.evaluate_adjoint(cur.reshape(-1), obs_ref.reshape(-1))
            assert loss == ref[0], (tile, tb)
            assert np.array_equal(grad, ref[1]), (tile, tb)


def test_streamed_xcd_tile_order_matches_oracle(orc):
    """A 600 x 200 grid (644 x 244 padded): 10 x 15 tiles of 64 x 16 a unit,
    enough for the launchers' XCD-aware tile order (from 128 tiles a unit).
    Gathers and gradient equal the oracle; the two-step adjoint and the
    two-step forward equal the single-step ones bit for bit."""
    from wgnet_amd.wave_gauss_rs import PyWaveEngine

    nt, nvx, nvz = 200, 600, 200
    E = PyWaveEngine(2, nt, 0.0008, 15.0, 20, 1000.0, 5000.0, nvx=nvx, nvz=nvz)
    assert E.streamed
    E.set_forward_time_block(1)
    E.set_time_block(1)
    nx, nz, _ = E.dims()
    assert ((nx - 4 + 63) // 64) * ((nz - 4 + 15) // 16) >= 128
    E.set_sources(np.array([[1000.0, 10.0], [4321.0, 10.0]], np.float32))
    true = layered(nvz, nvx, 5)
    cur = true.copy()
    cur[3:15, 80:520] += 120.0  # shallow enough for the 160 ms recorded
    obs = E.simulate(true.reshape(-1))
    for s, (x, z) in enumerate(E.sources()):
        g = orc.Geometry(nvx=nvx, nvz=nvz, nt=nt, dt=0.0008, src_x=float(x), src_z=float(z))
        err = relerr(obs[s], orc.forward_f32(g, true, E.wavelet()))
        print(f"600x200 shot {s}: gather rel-L2 {err:.3e}", flush=True)
        assert err <= 1e-5
    loss, grad = E.evaluate_adjoint(cur.reshape(-1), obs.reshape(-1))
    rl, rg = orc.evaluate_adjoint_f32(orc.Geometry(nvx=nvx, nvz=nvz, nt=nt, dt=0.0008), E.sources(), cur,
                                      E.wavelet(), obs)
    print(f"600x200: loss {loss:.9e} vs {rl:.9e}; grad rel-L2 {relerr(grad, rg):.3e}", flush=True)
    assert rl > 0.0 and np.linalg.norm(rg) > 0.0
    assert abs(loss - rl) <= 1e-6 * rl
    assert relerr(grad, rg) <= 1e-6
    E.set_time_block(2)
    loss2, grad2 = E.evaluate_adjoint(cur.reshape(-1), obs.reshape(-1))
    assert loss2 == loss and np.array_equal(grad2, grad)
    E.set_forward_time_block(2)
    assert np.array_equal(E.simulate(true.reshape(-1)), obs)
    loss3, grad3 = E.evaluate_adjoint(cur.reshape(-1), obs.reshape(-1))
    assert loss3 == loss and np.array_equal(grad3, grad)


def test_streamed_sources_inside_lean_rectangle_match_oracle(orc):
    """Sources deep inside the lean rectangle of the framed launches (a
    300 x 100 grid, 344 x 144 padded, wide enough for the aligned row
    layout): the lean kernels inject the source and add its gradient term
    themselves.  Gathers, loss and gradient equal the oracle; two steps per
    launch equal single steps bit for bit."""
    from wgnet_amd.wave_gauss_rs import PyWaveEngine

    nt, nvx, nvz = 400, 300, 100
    E = PyWaveEngine(2, nt, 0.0008, 15.0, 20, 1000.0, 5000.0, nvx=nvx, nvz=nvz)
    assert E.streamed
    E.set_band_split(False)
    E.set_forward_time_block(1)
    # cells (150, 12) and about (234.6, 15.3): inner cells (170, 32) and
    # (254.6, 35.3), inside the rectangle from inner cell (26, 26)
    E.set_sources(np.array([[1500.0, 120.0], [2345.6, 153.3]], np.float32))
    true = layered(nvz, nvx, 7)
    cur = true.copy()
    cur[8:20, 120:240] += 150.0
    obs = E.simulate(true.reshape(-1))
    for s, (x, z) in enumerate(E.sources()):
        g = orc.Geometry(nvx=nvx, nvz=nvz, nt=nt, dt=0.0008, src_x=float(x), src_z=float(z))
        err = relerr(obs[s], orc.forward_f32(g, true, E.wavelet()))
        print(f"300x100 deep shot {s}: gather rel-L2 {err:.3e}", flush=True)
        assert err <= 1e-5
    loss, grad = E.evaluate_adjoint(cur.reshape(-1), obs.reshape(-1))
    rl, rg = orc.evaluate_adjoint_f32(orc.Geometry(nvx=nvx, nvz=nvz, nt=nt, dt=0.0008), E.sources(), cur,
                                      E.wavelet(), obs)
    print(f"300x100 deep: loss {loss:.9e} vs {rl:.9e}; grad rel-L2 {relerr(grad, rg):.3e}", flush=True)
    assert rl > 0.0 and np.linalg.norm(rg) > 0.0
    assert abs(loss - rl) <= 1e-6 * rl
    assert relerr(grad, rg) <= 1e-6
    E.set_time_block(1)
    loss1, grad1 = E.evaluate_adjoint(cur.reshape(-1), obs.reshape(-1))
    assert loss1 == loss and np.array_equal(grad1, grad)
    E.set_time_block(2)
    E.set_forward_time_block(2)
    assert np.array_equal(E.simulate(true.reshape(-1)), obs)
    loss2, grad2 = E.evaluate_adjoint(cur.reshape(-1), obs.reshape(-1))
    assert loss2 == loss and np.array_equal(grad2, grad)
