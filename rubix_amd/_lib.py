"""ctypes binding of librubix_amd.so (the C ABI declared in include/rubix_amd.h).

torch is imported before the library is loaded so that the HIP runtime torch
brings (same SONAME libamdhip64.so.7) is the one the library binds to: one
runtime per process, so torch's streams and device pointers are valid here.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_float, c_int, c_int32, c_int64, c_size_t, c_uint64, c_void_p
from pathlib import Path

import torch  # noqa: F401  (load order, see module docstring)

LIB_PATH = Path(__file__).resolve().parent / "librubix_amd.so"


class RbxSsp(ctypes.Structure):
    _fields_ = [
        ("flux", c_void_p), ("ax_met", c_void_p), ("ax_age", c_void_p),
        ("wave_obs", c_void_p), ("cum", c_void_p), ("locate", c_void_p),
        ("n_met", c_int32), ("n_age", c_int32), ("n_wave", c_int32), ("n_locate", c_int32),
        ("locate_x0", c_float), ("locate_inv_h", c_float),
        ("log_axes", c_int32), ("locate_multi", c_int32),
    ]


class RbxTarget(ctypes.Structure):
    _fields_ = [("wave", c_void_p), ("dwave", c_void_p), ("n", c_int32),
                ("wmin", c_float), ("wmax", c_float), ("ascending", c_int32),
                ("pairs", c_void_p)]


class RbxStars(ctypes.Structure):
    _fields_ = [
        ("n", c_int64), ("mass", c_void_p), ("age", c_void_p), ("metallicity", c_void_p),
        ("velocity", c_void_p), ("vel_stride", c_int64), ("vel_comp", c_int32),
        ("n_spaxel", c_int32), ("n_keys", c_int32), ("reserved", c_int32),
        ("ids", c_void_p), ("speed_of_light", c_float), ("reserved_f", c_float),
    ]


class RbxGalaxy(ctypes.Structure):
    _fields_ = [("star_begin", c_int64), ("spaxel_begin", c_int32), ("nb", c_int32)]


class RbxTuning(ctypes.Structure):
    _fields_ = [("chunk", c_int32), ("waves", c_int32), ("persist", c_int32),
                ("flags", c_int32)]


class RbxRotation(ctypes.Structure):
    _fields_ = [("first", c_float * 9), ("second", c_float * 9)]


class RbxExtinction(ctypes.Structure):
    _fields_ = [("av", c_void_p), ("axav", c_void_p)]


class RbxNoiseStats(ctypes.Structure):
    _fields_ = [("median", c_float), ("median_lo", c_int32), ("median_hi", c_int32),
                ("any_nonpositive", c_int32)]


# name -> (restype, argtypes); mirrors include/rubix_amd.h
SIGNATURES = {
    "rbx_version": (c_int, []),
    "rbx_last_error_string": (ctypes.c_char_p, []),
    "rbx_spaxel_assign": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_int32, c_void_p, c_void_p]),
    "rbx_spaxel_sort_workspace_size": (c_size_t, [c_int64, c_int32]),
    "rbx_spaxel_sort": (c_int, [c_void_p, c_int64, c_int32, c_void_p, c_void_p, c_void_p,
                                c_void_p, c_size_t, c_void_p]),
    "rbx_ssp_prepare": (c_int, [POINTER(RbxSsp), c_void_p, c_void_p, c_void_p]),
    "rbx_cube_workspace_size": (c_size_t, [c_int64, c_int32, c_int32]),
    "rbx_cube_fwd": (c_int, [POINTER(RbxSsp), POINTER(RbxTarget), POINTER(RbxStars), c_void_p,
                             c_void_p, c_size_t, c_void_p]),
    "rbx_cube_bwd": (c_int, [POINTER(RbxSsp), POINTER(RbxTarget), POINTER(RbxStars), c_void_p,
                             c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_size_t, c_void_p]),
    "rbx_cube_fwd_ext": (c_int, [POINTER(RbxSsp), POINTER(RbxTarget), POINTER(RbxStars),
                                 POINTER(RbxExtinction), c_void_p, c_void_p, c_size_t, c_void_p]),
    "rbx_cube_bwd_ext": (c_int, [POINTER(RbxSsp), POINTER(RbxTarget), POINTER(RbxStars),
                                 POINTER(RbxExtinction), c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_size_t, c_void_p]),
    "rbx_batch_workspace_size": (c_size_t, [c_int64, c_int32, c_int32, POINTER(RbxTuning)]),
    "rbx_batch_fwd": (c_int, [POINTER(RbxSsp), POINTER(RbxTarget), POINTER(RbxStars), c_void_p,
                              c_int32, POINTER(RbxTuning), c_void_p, c_void_p, c_size_t,
                              c_void_p]),
    "rbx_batch_bwd": (c_int, [POINTER(RbxSsp), POINTER(RbxTarget), POINTER(RbxStars), c_void_p,
                              c_int32, POINTER(RbxTuning), c_void_p, c_void_p, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_size_t, c_void_p]),
    "rbx_conv_batch": (c_int, [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_int32,
                               c_void_p, c_int32, c_int32, c_void_p]),
    "rbx_conv_batch_dot": (c_int, [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32,
                                   c_int32, c_void_p, c_int32, c_int32, c_void_p, c_void_p,
                                   c_size_t, c_void_p]),
    "rbx_filter_particles": (c_int, [c_void_p, c_int64, c_int64, c_float, c_float, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_void_p]),
    "rbx_apply_mask": (c_int, [c_void_p, c_int64, c_int64, c_void_p, c_void_p, c_void_p]),
    "rbx_inertia_workspace_size": (c_size_t, [c_int64]),
    "rbx_inertia_tensor": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_float, c_void_p,
                                   c_void_p, c_size_t, c_void_p]),
    "rbx_rotate_rows": (c_int, [c_void_p, c_int64, c_int64, POINTER(RbxRotation), c_void_p,
                                c_void_p]),
    "rbx_noise_sample": (c_int, [c_int64, c_uint64, c_int32, c_void_p, c_void_p]),
    "rbx_noise_workspace_size": (c_size_t, [c_int32]),
    "rbx_apply_noise": (c_int, [c_void_p, c_int32, c_int32, c_void_p, c_uint64, c_int32, c_void_p,
                                c_void_p, c_void_p, c_void_p, c_void_p]),
    "rbx_apply_noise_bwd": (c_int, [c_void_p, c_void_p, c_int32, c_int32, c_void_p, c_void_p,
                                    c_void_p, c_uint64, c_int32, c_void_p, c_size_t, c_void_p,
                                    c_void_p]),
    "rbx_extinction_curve": (c_int, [c_int32, c_float, c_void_p, c_int32, c_void_p, c_void_p]),
    "rbx_gas_extinction": (c_int, [c_void_p, c_void_p, c_int64, c_int64, c_int32, c_int32, c_float,
                                   c_float, c_void_p, c_void_p]),
    "rbx_column_av_workspace_size": (c_size_t, [c_int64, c_int32]),
    "rbx_column_av": (c_int, [c_void_p, c_int64, c_void_p, c_void_p, c_int64, c_void_p, c_int64,
                              c_void_p, c_int64, c_int32, c_int32, c_void_p, c_size_t, c_void_p,
                              c_void_p]),
    "rbx_apply_extinction": (c_int, [c_void_p, c_void_p, c_void_p, c_int64, c_int32, c_void_p,
                                     c_void_p]),
    "rbx_ssp_lookup": (c_int, [POINTER(RbxSsp), c_void_p, c_void_p, c_int64, c_void_p, c_void_p]),
    "rbx_ssp_lookup_bwd": (c_int, [POINTER(RbxSsp), c_void_p, c_void_p, c_int64, c_void_p, c_void_p,
                                   c_void_p, c_void_p]),
    "rbx_scale_rows": (c_int, [c_void_p, c_void_p, c_int64, c_int32, c_void_p, c_void_p]),
    "rbx_row_dot": (c_int, [c_void_p, c_void_p, c_int64, c_int32, c_void_p, c_void_p]),
    "rbx_resample": (c_int, [POINTER(RbxSsp), POINTER(RbxTarget), c_void_p, c_void_p, c_int64,
                             c_int32, c_float, c_int64, c_void_p, c_void_p]),
    "rbx_resample_bwd": (c_int, [POINTER(RbxSsp), POINTER(RbxTarget), c_void_p, c_void_p, c_int64,
                                 c_int32, c_float, c_int64, c_void_p, c_void_p, c_void_p, c_void_p]),
    "rbx_segment_cube_workspace_size": (c_size_t, [c_int64, c_int32, c_int32]),
    "rbx_segment_cube": (c_int, [c_void_p, c_void_p, c_int64, c_int32, c_int32, c_int32, c_void_p,
                                 c_void_p, c_size_t, c_void_p]),
    "rbx_segment_cube_bwd": (c_int, [c_void_p, c_void_p, c_int64, c_int32, c_int32, c_void_p,
                                     c_void_p]),
    "rbx_conv_axis": (c_int, [c_void_p, c_void_p, c_int64, c_int32, c_int64, c_void_p, c_int32,
                              c_int32, c_void_p]),
    "rbx_psf_sep": (c_int, [c_void_p, c_void_p, c_void_p, c_int32, c_int32, c_int32, c_void_p,
                            c_int32, c_int32, c_void_p, c_int32, c_int32, c_void_p]),
    "rbx_conv2d_planes": (c_int, [c_void_p, c_void_p, c_int32, c_int32, c_int32, c_void_p, c_int32,
                                  c_int32, c_int32, c_int32, c_void_p]),
    "rbx_dot_workspace_size": (c_size_t, [c_int64]),
    "rbx_conv_axis_dot": (c_int, [c_void_p, c_void_p, c_int64, c_int32, c_int64, c_void_p, c_int32,
                                  c_int32, c_void_p, c_void_p, c_size_t, c_void_p]),
    "rbx_conv_axis_dot_noalloc": (c_int, [c_void_p, c_void_p, c_int64, c_int32, c_int64, c_void_p,
                                          c_int32, c_int32, c_void_p, c_void_p]),
}

_lib = None


def load():
    """Load and declare the library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH.name} is not built; run `python -m rubix_amd.build` "
                "(hipcc, gfx950) first")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class RbxError(RuntimeError):
    """A rubix_amd C-ABI call returned a non-zero status."""

    def __init__(self, what, code, msg):
        super().__init__(f"{what} failed (status {code}): {msg}")
        self.code = code


def check(rc, what):
    if rc != 0:
        msg = load().rbx_last_error_string()
        raise RbxError(what, rc, msg.decode() if msg else "")


class CallTimer:
    """Optional per-entry-point GPU timing: when enabled, every status call is
    bracketed by HIP events on torch's current stream (the stream the call
    launches on), so the elapsed time covers exactly that call's kernels."""

    def __init__(self):
        self.enabled = False
        self.events = {}

    def start(self):
        self.enabled = True
        self.events = {}

    def stop(self):
        """Synchronise and return {name: [ms, ...]}."""
        self.enabled = False
        torch.cuda.synchronize()
        return {k: [a.elapsed_time(b) for a, b in v] for k, v in self.events.items()}


TIMER = CallTimer()


def call(name, *args):
    """Call a status-returning entry point and raise RbxError on failure."""
    if TIMER.enabled:
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        rc = getattr(load(), name)(*args)
        b.record()
        TIMER.events.setdefault(name, []).append((a, b))
    else:
        rc = getattr(load(), name)(*args)
    check(rc, name)


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return c_void_p(t.data_ptr())


def stream_handle(device=None):
    return c_void_p(torch.cuda.current_stream(device).cuda_stream)

