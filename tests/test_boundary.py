"""The C-ABI boundary on the CPU: the library loads, exports every symbol the
header declares, its host-side functions equal the oracle bit for bit, and
argument checks raise the reference's exception types before any device
work (crates/wave-gauss-rs/src/lib.rs:33-35, 84-86, 100-103, 110-113)."""

from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols() -> list[str]:
    text = open(os.path.join(ROOT, "include", "wgnet_hip.h")).read()
    return sorted(set(re.findall(r"\b(wg_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_header_symbols():
    from wgnet_amd import _lib

    lib = _lib.load()
    syms = header_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_config_defaults_match_reference():
    from wgnet_amd import _lib

    c = _lib.Config()
    _lib.load().wg_config_default(ctypes.byref(c))
    # lib.rs:52 and geology-core lib.rs:5-8
    assert (c.shots, c.nt, c.pml, c.nvx, c.nvz) == (5, 2001, 20, 70, 70)
    assert (c.dt, c.peak_hz, c.vmin, c.vmax, c.h) == (np.float32(0.0005), 15.0, 1000.0, 5000.0, 10.0)


def test_host_ricker_and_cpml_equal_oracle(orc):
    from wgnet_amd import wave_gauss_rs as W
    from wgnet_amd import _lib

    for n, dt, f in [(2001, 0.0005, 15.0), (1000, 0.001, 15.0), (4000, 0.0008, 9.0)]:
        np.testing.assert_array_equal(W.ricker_wavelet(n, dt, f), orc.ricker_f32(n, dt, f))
    lib = _lib.load()
    for n in (114, 144, 185, 54):
        a, b, db = (np.zeros(n, np.float32) for _ in range(3))
        lib.wg_cpml_profiles(n, 10.0, 0.001, 20, 5000.0, 15.0, ctypes.c_void_p(a.ctypes.data),
                             ctypes.c_void_p(b.ctypes.data), ctypes.c_void_p(db.ctypes.data))
        ra, rb, rdb = orc.cpml_f32(n, 10.0, 0.001, 20, 5000.0, 15.0)
        np.testing.assert_array_equal(a, ra)
        np.testing.assert_array_equal(b, rb)
        np.testing.assert_array_equal(db, rdb)


def test_invalid_config_rejected_before_device():
    from wgnet_amd import _lib

    lib = _lib.load()
    c = _lib.Config()
    lib.wg_config_default(ctypes.byref(c))
    c.vmin, c.vmax = 5000.0, 1000.0
    h = ctypes.c_void_p()
    assert lib.wg_engine_create(ctypes.byref(c), ctypes.byref(h)) == _lib.WG_E_INVALID
    assert b"vmin" in lib.wg_last_error()
    with pytest.raises(ValueError):
        _lib.check(_lib.WG_E_INVALID)


def test_shot_solver_argument_checks():
    from wgnet_amd import wave_gauss_rs as W

    with pytest.raises(TypeError):
        W.simulate_shot_cpu(np.zeros(4900, np.float64), 10.0, 0.001, 10, 20, 1000.0, 5000.0, 15.0, 0.0, 10.0)
    with pytest.raises(ValueError, match="4900"):
        W.simulate_shot_cpu(np.zeros(4901, np.float32), 10.0, 0.001, 10, 20, 1000.0, 5000.0, 15.0, 0.0, 10.0)
    with pytest.raises(ValueError, match="contiguous"):
        W.simulate_shot_cpu(np.zeros(9800, np.float32)[::2], 10.0, 0.001, 10, 20, 1000.0, 5000.0, 15.0, 0.0, 10.0)


def test_lowpass_matches_oracle_restatement():
    from oracle import bandlimit as BL
    from wgnet_amd.physics import lowpass_kernel

    for f in (6.0, 8.0, 9.0, 12.0, 15.0):
        for dt in (0.001, 0.0005, 0.0008):
            np.testing.assert_array_equal(lowpass_kernel(f, dt), BL.lowpass_kernel(f, dt))


def test_product_does_not_import_oracle():
    pkg = os.path.join(ROOT, "wgnet_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".h", ".cpp")):
                text = open(os.path.join(dirpath, fn)).read()
                assert "oracle" not in re.sub(r"#.*|//.*", "", text).lower().replace("oracle-free", ""), fn
