"""ctypes binding of libwgnet_hip.so (the C-ABI in include/wgnet_hip.h).

torch is imported first so that the library binds to the HIP runtime torch
already loaded (both use the soname libamdhip64.so.7); device pointers from
torch tensors and torch streams are then valid in the library.

There is no fallback: if the library is missing or fails to load, every entry
point raises.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwgnet_hip.so")

WG_OK = 0
WG_E_INVALID = -1
WG_E_DEVICE = -2
WG_E_TOO_LARGE = -3
WG_E_ZERO_ENERGY = -4


class Config(ctypes.Structure):
    _fields_ = [
        ("shots", ctypes.c_int32),
        ("nt", ctypes.c_int32),
        ("dt", ctypes.c_float),
        ("peak_hz", ctypes.c_float),
        ("pml", ctypes.c_int32),
        ("vmin", ctypes.c_float),
        ("vmax", ctypes.c_float),
        ("nvx", ctypes.c_int32),
        ("nvz", ctypes.c_int32),
        ("h", ctypes.c_float),
    ]


_f32p = ctypes.POINTER(ctypes.c_float)
_vp = ctypes.c_void_p
_i32 = ctypes.c_int32

# name -> (restype, argtypes); the complete exported surface of wgnet_hip.h.
SIGNATURES = {
    "wg_config_default": (None, [ctypes.POINTER(Config)]),
    "wg_last_error": (ctypes.c_char_p, []),
    "wg_engine_create": (_i32, [ctypes.POINTER(Config), ctypes.POINTER(_vp)]),
    "wg_engine_destroy": (None, [_vp]),
    "wg_engine_shots": (_i32, [_vp]),
    "wg_engine_nt": (_i32, [_vp]),
    "wg_engine_nrec": (_i32, [_vp]),
    "wg_engine_dims": (_i32, [_vp, ctypes.POINTER(_i32), ctypes.POINTER(_i32), ctypes.POINTER(_i32)]),
    "wg_get_sources": (_i32, [_vp, _vp, _i32]),
    "wg_set_sources": (_i32, [_vp, _vp, _i32]),
    "wg_get_wavelet": (_i32, [_vp, _vp, _i32]),
    "wg_set_wavelet": (_i32, [_vp, _vp, _i32]),
    "wg_simulate": (_i32, [_vp, _vp, _vp]),
    "wg_evaluate_adjoint": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "wg_simulate_dev": (_i32, [_vp, _i32, _vp, _vp, _vp]),
    "wg_loss_grad_dev": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "wg_loss_grad_true_dev": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "wg_loss_grad_wavelets_dev": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "wg_loss_dev": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "wg_loss_noalloc_dev": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "wg_source_corners": (_i32, [_vp, _i32, _vp, _vp, _vp]),
    "wg_loss_grad_units_dev": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "wg_loss_grad_groups_dev": (_i32, [_i32, _vp, _vp]),
    "wg_simulate_units_dev": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "wg_adjoint_apply_dev": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "wg_born_dev": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp]),
    "wg_simulate_wavelets_dev": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp]),
    "wg_set_snapshot_budget": (_i32, [_vp, ctypes.c_uint64]),
    "wg_engine_streamed": (_i32, [_vp]),
    "wg_stream_tiles": (_i32, []),
    "wg_engine_bands": (_i32, [_vp]),
    "wg_set_band_split": (_i32, [_vp, _i32]),
    "wg_illumination_dev": (_i32, [_vp, _i32, _vp, _vp, _vp]),
    "wg_set_overlap": (_i32, [_vp, _i32]),
    "wg_set_time_block": (_i32, [_vp, _i32]),
    "wg_set_forward_time_block": (_i32, [_vp, _i32]),
    "wg_set_time_split": (_i32, [_vp, _i32]),
    "wg_set_queue_wait_limit": (_i32, [_vp, _i32]),
    "wg_causal_filter_dev": (_i32, [_vp, ctypes.c_int64, _i32, _i32, _vp, _i32, _vp, _vp]),
    "wg_set_stream_tile": (_i32, [_vp, _i32]),
    "wg_timing_enable": (_i32, [_vp, _i32]),
    "wg_timing_read": (_i32, [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]),
    "wg_ricker_wavelet": (None, [_i32, ctypes.c_float, ctypes.c_float, _vp]),
    "wg_cpml_profiles": (None, [_i32, ctypes.c_float, ctypes.c_float, _i32, ctypes.c_float, ctypes.c_float, _vp, _vp, _vp]),
    "wg_pc_last_error": (ctypes.c_char_p, []),
    "wg_pc_create": (_i32, [_i32, _i32, _i32, ctypes.c_uint64, ctypes.POINTER(_vp)]),
    "wg_pc_destroy": (None, [_vp]),
    "wg_pc_num_params": (_i32, []),
    "wg_pc_num_layers": (_i32, []),
    "wg_pc_layer_shape": (_i32, [_i32, _vp, _vp, _vp]),
    "wg_pc_get_params": (_i32, [_vp, _vp, _i32]),
    "wg_pc_set_params": (_i32, [_vp, _vp, _i32]),
    "wg_pc_forward_dev": (_i32, [_vp, _i32, _vp, _vp, _vp]),
    "wg_pc_loss_grad": (_i32, [_vp, _i32, _vp, _vp, _vp, _vp, _vp]),
    "wg_pc_train_step_dev": (_i32, [_vp, _i32, _vp, _vp, ctypes.c_float, ctypes.c_float, _vp, _vp]),
}

_LIB: ctypes.CDLL | None = None


class WaveEngineError(RuntimeError):
    """HIP runtime failure inside the wave engine."""


class GroupArgs(ctypes.Structure):
    """wg_group_args of include/wgnet_hip.h: one group of wg_loss_grad_groups_dev."""

    _fields_ = [("engine", _vp), ("B", _i32), ("src_cell_dev", _vp), ("src_w_dev", _vp), ("active_dev", _vp),
                ("vel_dev", _vp), ("wav_dev", _vp), ("obs_dev", _vp), ("loss_dev", _vp), ("grad_dev", _vp),
                ("traces_dev", _vp), ("energy_dev", _vp)]


def load() -> ctypes.CDLL:
    """Loads libwgnet_hip.so (raises if it is missing; build with
    `python -m wgnet_amd.build`)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    import torch  # noqa: F401  (binds the library to torch's HIP runtime)

    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m wgnet_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(rc: int, pc: bool = False) -> None:
    """Maps a status code to the reference's Python exception types (`pc`:
    the status came from a wg_pc_* call)."""
    if rc == WG_OK:
        return
    lib = load()
    msg = (lib.wg_pc_last_error() if pc else lib.wg_last_error()).decode("utf-8", "replace")
    if rc in (WG_E_INVALID, WG_E_TOO_LARGE, WG_E_ZERO_ENERGY):
        raise ValueError(msg)
    raise WaveEngineError(msg)
