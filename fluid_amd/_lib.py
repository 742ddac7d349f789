"""Loader of the in-tree HIP library libfluidhip.so (C ABI: include/fluid_hip.h).

There is no CPU fallback: if the library is missing or no GPU is present the calls
fail loudly.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("FLUID_AMD_LIB", os.path.join(HERE, "libfluidhip.so"))

FL_OK = 0
PARTICLES_NOALLOC_MAX = 131072  # FL_PARTICLES_NOALLOC_MAX
STATUS = {0: "FL_OK", 1: "FL_ERR_INVALID_ARG", 2: "FL_ERR_X_RANGE", 3: "FL_ERR_Y_RANGE", 4: "FL_ERR_NO_DEVICE",
          5: "FL_ERR_HIP", 6: "FL_ERR_UNSUPPORTED", 7: "FL_ERR_INTERNAL"}


class FlParams(ctypes.Structure):
    _fields_ = [
        ("confinement", ctypes.c_float),
        ("viscosity_diffusion", ctypes.c_float),
        ("pressure_damping", ctypes.c_float),
        ("turbulence_strength", ctypes.c_float),
        ("smoke_advection", ctypes.c_float),
        ("relaxation", ctypes.c_float),
        ("use_multigrid", ctypes.c_int32),
        ("multigrid_levels", ctypes.c_int32),
        ("use_bfecc", ctypes.c_int32),
        ("pressure_iters", ctypes.c_int32),
        ("early_exit", ctypes.c_int32),
        ("solver", ctypes.c_int32),
        ("block_sweeps", ctypes.c_int32),
        ("precise_advection", ctypes.c_int32),
        ("handoff_lag", ctypes.c_int32),
        ("block_cols", ctypes.c_int32),
    ]


class FlStats(ctypes.Structure):
    _fields_ = [
        ("sweeps_run", ctypes.c_int32),
        ("reran", ctypes.c_int32),
        ("invalid_sample", ctypes.c_int32),
        ("max_div_last", ctypes.c_float),
    ]


class FlGsLayout(ctypes.Structure):
    _fields_ = [
        ("sweeps_per_tile", ctypes.c_int32),
        ("cols_per_block", ctypes.c_int32),
        ("handoff_lag", ctypes.c_int32),
        ("row_tiles", ctypes.c_int32),
        ("sweep_blocks", ctypes.c_int32),
        ("steps_per_tile", ctypes.c_int32),
        ("tiles", ctypes.c_int32),
        ("workgroups", ctypes.c_int32),
    ]


class FlTiming(ctypes.Structure):
    _fields_ = [
        ("total_ms", ctypes.c_double),
        ("pressure_ms", ctypes.c_double),
        ("advect_ms", ctypes.c_double),
        ("other_ms", ctypes.c_double),
        ("steps", ctypes.c_int32),
        ("pad", ctypes.c_int32),
    ]


# Every symbol declared in include/fluid_hip.h with its ctypes signature.
_vp, _i, _f, _sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_size_t
_ip, _fp = ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_float)
SIGNATURES = {
    "fl_last_error": (ctypes.c_char_p, []),
    "fl_ctx_last_error": (_i, [_vp, ctypes.c_char_p, _sz]),
    "fl_device_count": (_i, [_ip]),
    "fl_create": (_i, [_f, _i, _i, _f, ctypes.POINTER(_vp)]),
    "fl_destroy": (_i, [_vp]),
    "fl_dims": (_i, [_vp, _ip, _ip]),
    "fl_get_h": (_i, [_vp, _fp]),
    "fl_get_params": (_i, [_vp, ctypes.POINTER(FlParams)]),
    "fl_set_params": (_i, [_vp, ctypes.POINTER(FlParams)]),
    "fl_upload": (_i, [_vp, _i, _vp, _sz]),
    "fl_download": (_i, [_vp, _i, _vp, _sz]),
    "fl_fill_cells": (_i, [_vp, _i, _i, _i, _i, _i, _f]),
    "fl_simulate": (_i, [_vp, _f]),
    "fl_sync": (_i, [_vp]),
    "fl_get_stats": (_i, [_vp, ctypes.POINTER(FlStats)]),
    "fl_sweep_maxdiv": (_i, [_vp, _vp, _i, _ip]),
    "fl_get_gs_layout": (_i, [_vp, ctypes.POINTER(FlGsLayout)]),
    "fl_set_timing": (_i, [_vp, _i]),
    "fl_get_timing": (_i, [_vp, ctypes.POINTER(FlTiming)]),
    "fl_set_solid": (_i, [_vp, _i, _i, _i]),
    "fl_is_solid": (_i, [_vp, _i, _i, _ip]),
    "fl_set_velocity": (_i, [_vp, _i, _i, _f, _f]),
    "fl_add_smoke": (_i, [_vp, _i, _i, _f]),
    "fl_reset": (_i, [_vp]),
    "fl_apply_force": (_i, [_vp, _i, _i, _f, _f]),
    "fl_apply_force_radius": (_i, [_vp, _i, _i, _f, _f, _i]),
    "fl_set_circular_obstacle": (_i, [_vp, _i, _i, _i]),
    "fl_get_adaptive_time_step": (_i, [_vp, _f, _fp]),
    "fl_max_divergence": (_i, [_vp, _fp]),
    "fl_sample_velocity": (_i, [_vp, _f, _f, _fp, _fp]),
    "fl_sample_velocity_batch": (_i, [_vp, _vp, _vp, _i, _vp, _vp]),
    "fl_vorticity": (_i, [_vp, _vp, _sz, _fp, _fp]),
    "fl_velocity_magnitude": (_i, [_vp, _vp, _sz, _fp, _fp]),
    "fl_field_minmax": (_i, [_vp, _i, _fp, _fp]),
    "fl_pass": (_i, [_vp, _i, _f, _i]),
    "fl_scene_set_walls": (_i, [_vp, _i, _i, _i, _i]),
    "fl_scene_apply_walls": (_i, [_vp]),
    "fl_scene_clear_walls": (_i, [_vp]),
    "fl_scene_preset": (_i, [_vp, _i, _f]),
    "fl_scene_set_jet": (_i, [_vp, _i, _i, _i, _i, _f, _f]),
    "fl_scene_add_source": (_i, [_vp, _i, _i, _i, _f, _f]),
    "fl_scene_clear_sources": (_i, [_vp]),
    "fl_scene_apply_sources": (_i, [_vp]),
    "fl_scene_update": (_i, [_vp, _f]),
    "fl_scene_draw_line": (_i, [_vp, _i, _i, _i, _i, _i]),
    "fl_scene_smoke_brush": (_i, [_vp, _i, _i, _i, _f]),
    "fl_apply_force_gaussian": (_i, [_vp, _i, _i, _f, _f, _f]),
    "fl_render": (_i, [_vp, _i, _i, _vp, _sz]),
    "fl_advect_particles": (_i, [_vp, _f, _i, _vp, _vp, _vp, _vp, _vp]),
    "fl_advect_particles_noalloc": (_i, [_vp, _f, _i, _vp, _vp, _vp, _vp, _vp]),
    "fl_autotune": (_i, [_vp, _f, _ip, ctypes.POINTER(_f)]),
    "fl_batch_create": (_i, [_i, _ip, _ip, _f, _f, ctypes.POINTER(_vp)]),
    "fl_batch_destroy": (_i, [_vp]),
    "fl_batch_size": (_i, [_vp, _ip]),
    "fl_batch_member": (_i, [_vp, _i, ctypes.POINTER(_vp)]),
    "fl_batch_get_params": (_i, [_vp, ctypes.POINTER(FlParams)]),
    "fl_batch_set_params": (_i, [_vp, ctypes.POINTER(FlParams)]),
    "fl_batch_update": (_i, [_vp, _f]),
    "fl_batch_sync": (_i, [_vp]),
    "fl_batch_set_graph": (_i, [_vp, _i]),
    "fl_batch_autotune": (_i, [_vp, _f, _ip, ctypes.POINTER(_f)]),
    "fl_vjp_create": (_i, [_vp, _i, _i, ctypes.POINTER(_vp)]),
    "fl_vjp_destroy": (_i, [_vp]),
    "fl_vjp_forward": (_i, [_vp, _f, _fp, _fp, _fp, _sz, ctypes.POINTER(ctypes.c_double)]),
    "fl_vjp_backward": (_i, [_vp, _fp, _fp, _sz]),
}

_lib = None


class FluidError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s: %s" % (STATUS.get(code, code), msg))
        self.code = code


def load():
    """Load libfluidhip.so (built by fluid_amd/build.py). Raises if it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                "libfluidhip.so is not built (run `python fluid_amd/build.py`); there is no CPU fallback")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != FL_OK:
        raise FluidError(rc, load().fl_last_error().decode())
    return rc
