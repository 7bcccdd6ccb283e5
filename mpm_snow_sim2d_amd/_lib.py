"""Loads libmpm_hip.so (built in-tree by mpm_snow_sim2d_amd.build) and declares the
C-ABI of include/mpm/mpm.h for ctypes.  There is no fallback: if the library is
missing the import of the solver fails with the build command to run."""
from __future__ import annotations

import ctypes
import os

from .params import COptions, Params

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpm_hip.so")

MPM_OK = 0
MPM_E_INVALID = -1
MPM_E_HIP = -2
MPM_E_STATE = -3
MPM_E_NONFINITE = -4

ADVANCE_PLAIN, ADVANCE_NETWORK, ADVANCE_BALLISTIC = 0, 1, 2
CONV_AUTO, CONV_STREAM, CONV_LDS = 0, 1, 2

KERNEL_P2G, KERNEL_GRID_UPDATE, KERNEL_G2P, KERNEL_RESET, KERNEL_G2P2G, KERNEL_SORT = range(6)

# Every exported symbol of include/mpm/mpm.h (checked by the CPU test suite).
EXPORTS = [
    "mpm_params_reference", "mpm_params_derive_lame", "mpm_options_default",
    "mpm_create", "mpm_destroy", "mpm_step",
    "mpm_p2g", "mpm_grid_update", "mpm_g2p", "mpm_reset_grid",
    "mpm_sort", "mpm_get_sort", "mpm_sync", "mpm_get_positions",
    "mpm_get_state", "mpm_set_state", "mpm_get_grid", "mpm_get_grid_force",
    "mpm_num_particles", "mpm_num_nodes",
    "mpm_set_timing", "mpm_set_timing_mask", "mpm_set_timing_every", "mpm_reset_timers", "mpm_get_kernel_time",
    "mpm_check_finite", "mpm_sort_pairs_host", "mpm_sort_pairs_host_noalloc", "mpm_last_error",
    "mpm_batch_create", "mpm_num_scenes", "mpm_scene_range", "mpm_set_graph", "mpm_autotune",
    "mpm_surrogate_create", "mpm_surrogate_destroy", "mpm_surrogate_num_params", "mpm_surrogate_get_params",
    "mpm_surrogate_set_params", "mpm_surrogate_set_scale", "mpm_surrogate_set_conv", "mpm_surrogate_set_macro", "mpm_surrogate_set_targets", "mpm_surrogate_forward", "mpm_surrogate_gradient",
    "mpm_surrogate_record", "mpm_surrogate_get_sample", "mpm_surrogate_train", "mpm_surrogate_baseline_loss",
    "mpm_surrogate_advance", "mpm_surrogate_rollout", "mpm_surrogate_time",
    "mpm_save_snapshot", "mpm_snapshot_info", "mpm_load_snapshot",
    "mpm_camera_default", "mpm_render_frame", "mpm_write_png",
]


class CParticles(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int64),
        ("x", ctypes.c_void_p),
        ("v", ctypes.c_void_p),
        ("mass", ctypes.c_void_p),
        ("vol", ctypes.c_void_p),
        ("C", ctypes.c_void_p),
        ("F", ctypes.c_void_p),
        ("Jp", ctypes.c_void_p),
    ]


class CCamera(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in
                ("angle_deg", "elevation_deg", "distance", "fov_deg", "z_near", "z_far")]


class MPMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mpm error {code}: {msg}")
        self.code = code


_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -m mpm_snow_sim2d_amd.build`")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    vp = ctypes.c_void_p
    i32, i64 = ctypes.c_int, ctypes.c_int64
    sig = {
        "mpm_params_reference": (i32, [i32, P(Params)]),
        "mpm_params_derive_lame": (i32, [P(Params)]),
        "mpm_options_default": (i32, [P(COptions)]),
        "mpm_create": (i32, [P(Params), P(CParticles), P(COptions), P(vp)]),
        "mpm_destroy": (i32, [vp]),
        "mpm_step": (i32, [vp, i32]),
        "mpm_p2g": (i32, [vp]),
        "mpm_grid_update": (i32, [vp]),
        "mpm_g2p": (i32, [vp]),
        "mpm_reset_grid": (i32, [vp]),
        "mpm_sort": (i32, [vp]),
        "mpm_get_sort": (i32, [vp, vp, vp]),
        "mpm_sync": (i32, [vp]),
        "mpm_get_positions": (i32, [vp, vp]),
        "mpm_get_state": (i32, [vp, vp, vp, vp, vp, vp]),
        "mpm_set_state": (i32, [vp, vp, vp, vp, vp, vp]),
        "mpm_get_grid": (i32, [vp, vp, vp, vp, vp]),
        "mpm_get_grid_force": (i32, [vp, vp]),
        "mpm_num_particles": (i32, [vp, P(i64)]),
        "mpm_num_nodes": (i32, [P(Params), P(i64)]),
        "mpm_set_timing": (i32, [vp, i32]),
        "mpm_set_timing_mask": (i32, [vp, ctypes.c_uint32]),
        "mpm_set_timing_every": (i32, [vp, i32]),
        "mpm_reset_timers": (i32, [vp]),
        "mpm_get_kernel_time": (i32, [vp, i32, P(ctypes.c_double), P(i64)]),
        "mpm_check_finite": (i32, [vp, P(i64)]),
        "mpm_sort_pairs_host": (i32, [vp, i64, i32, vp, vp]),
        "mpm_sort_pairs_host_noalloc": (i32, [vp, i64, i32, vp, vp]),
        "mpm_last_error": (ctypes.c_char_p, []),
        "mpm_batch_create": (i32, [i32, P(Params), P(CParticles), P(COptions), P(vp)]),
        "mpm_num_scenes": (i32, [vp, P(i32)]),
        "mpm_scene_range": (i32, [vp, i32, P(i64), P(i64)]),
        "mpm_set_graph": (i32, [vp, i32]),
        "mpm_autotune": (i32, [i32, P(Params), P(CParticles), i32, i32, P(COptions), P(ctypes.c_double)]),
        "mpm_surrogate_create": (i32, [i32, i32, i32, i32, ctypes.c_uint64, P(vp)]),
        "mpm_surrogate_destroy": (i32, [vp]),
        "mpm_surrogate_num_params": (i32, [vp, P(i64)]),
        "mpm_surrogate_get_params": (i32, [vp, vp]),
        "mpm_surrogate_set_params": (i32, [vp, vp]),
        "mpm_surrogate_set_scale": (i32, [vp, ctypes.c_float]),
        "mpm_surrogate_set_conv": (i32, [vp, i32]),
        "mpm_surrogate_set_macro": (i32, [vp, i32]),
        "mpm_surrogate_set_targets": (i32, [vp, i32]),
        "mpm_surrogate_forward": (i32, [vp, i32, vp, vp, vp]),
        "mpm_surrogate_gradient": (i32, [vp, i32, vp, vp, vp, vp, vp]),
        "mpm_surrogate_record": (i32, [vp, vp, i32, i32, i32, i32]),
        "mpm_surrogate_get_sample": (i32, [vp, i32, vp, vp]),
        "mpm_surrogate_train": (i32, [vp, vp, i32, ctypes.c_float, vp]),
        "mpm_surrogate_baseline_loss": (i32, [vp, vp, i32, vp]),
        "mpm_surrogate_advance": (i32, [vp, vp, i32, i32, ctypes.c_float, vp, vp]),
        "mpm_surrogate_rollout": (i32, [vp, vp, i32, i32, i32, ctypes.c_float, vp, vp, vp]),
        "mpm_surrogate_time": (i32, [vp, i32, i32, i32, P(ctypes.c_double)]),
        "mpm_save_snapshot": (i32, [vp, ctypes.c_char_p]),
        "mpm_snapshot_info": (i32, [ctypes.c_char_p, P(Params), P(i64), P(i64)]),
        "mpm_load_snapshot": (i32, [ctypes.c_char_p, P(COptions), P(vp)]),
        "mpm_camera_default": (i32, [P(CCamera)]),
        "mpm_render_frame": (i32, [vp, P(CCamera), i32, i32, i32, vp]),
        "mpm_write_png": (i32, [ctypes.c_char_p, i32, i32, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(code: int) -> int:
    if code != MPM_OK:
        msg = load().mpm_last_error()
        raise MPMError(code, msg.decode() if msg else "")
    return code
