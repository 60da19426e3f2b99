"""ctypes binding of libhdns.so, the C ABI declared in include/hdns.h.

This is the same binding a Julia ``ccall`` shim makes; it loads the in-tree library and fails
loudly when it has not been built -- there is no CPU fallback behind it.
"""
from __future__ import annotations

import ctypes
import re
from ctypes import POINTER, c_char_p, c_double, c_int32, c_int64, c_uint8, c_void_p
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIBPATH = HERE / "libhdns.so"
HEADER = HERE.parent / "include" / "hdns.h"

OK, ERR_INVALID_ARGUMENT, ERR_STATE, ERR_NONFINITE, ERR_DEVICE, ERR_UNSUPPORTED = range(6)
FP32_BULK, FP64 = 0, 1
RK3, RK4, SSPRK3, HEUN = 3, 4, 5, 6
NO_FORCING, CONSTANT_POWER, LINEAR, VARIANCE_CONTROLLER, VARIANCE_RESCALING, STOCHASTIC_OU = 0, 1, 2, 3, 4, 5
OVER_BAND, OVER_TOTAL = 0, 1
P_HEUN, P_ABM, P_AB = 0, 1, 2
FIELD_VELOCITY, FIELD_SCALAR, FIELD_VORTICITY = 0, 1, 2


class HdnsConfig(ctypes.Structure):
    _fields_ = [
        ("device", c_int32), ("grid_points", c_int32), ("side_length", c_double),
        ("precision", c_int32), ("scheme", c_int32), ("dealiasing", c_int32), ("has_scalar", c_int32),
        ("viscosity", c_double), ("schmidt", c_double),
        ("velocity_forcing", c_int32), ("injection_rate", c_double),
        ("lowest_shell", c_int32), ("highest_shell", c_int32),
        ("mean_gradient", c_double), ("tracer_capacity", c_int64),
        ("interpolation", c_int32), ("particle_scheme", c_int32),
        ("store_physical", c_int32), ("forcing_over", c_int32),
        ("forcing_target", c_double), ("forcing_gain", c_double), ("forcing_time", c_double),
        ("particle_steps", c_int32), ("unwrapped", c_int32),
        ("scalar_forcing", c_int32), ("scalar_lowest_shell", c_int32), ("scalar_highest_shell", c_int32),
        ("scalar_over", c_int32), ("scalar_rate", c_double), ("scalar_target", c_double),
        ("scalar_gain", c_double), ("scalar_time", c_double),
        ("filter_order", c_int32), ("filter_strength", c_double),
    ]


class HdnsError(RuntimeError):
    """A failure the library reported (device errors, schedule errors)."""


class BlowUpError(HdnsError):
    """A non-finite velocity or scalar after the closing deposit (src/SimulationRun.jl:1216-1224)."""


_lib = None

_D = POINTER(c_double)
_SIGS = {
    "hdns_ctx_create": [POINTER(HdnsConfig), POINTER(c_void_p)],
    "hdns_ctx_destroy": [c_void_p],
    "hdns_memory_bytes": [c_void_p, POINTER(c_int64)],
    "hdns_state_upload": [c_void_p, _D, _D],
    "hdns_state_download": [c_void_p, _D, _D],
    "hdns_step": [c_void_p, c_double],
    "hdns_step_graph": [c_void_p, c_double],
    "hdns_deposit": [c_void_p],
    "hdns_max_speed": [c_void_p, _D],
    "hdns_reductions": [c_void_p, _D],
    "hdns_spectrum": [c_void_p, c_int32, _D, c_int32],
    "hdns_spectrum_noalloc": [c_void_p, c_int32, _D, c_int32],
    "hdns_state_distance": [c_void_p, c_void_p, _D],
    "hdns_derivative_skewness": [c_void_p, _D],
    "hdns_field_download": [c_void_p, c_int32, _D],
    "hdns_tracers_upload": [c_void_p, c_int64, _D],
    "hdns_tracers_download": [c_void_p, _D],
    "hdns_tracers_unwrapped_download": [c_void_p, _D],
    "hdns_tracers_unwrapped_upload": [c_void_p, c_int64, _D],
    "hdns_rebuild_traversal": [c_void_p],
    "hdns_advect": [c_void_p, c_double],
    "hdns_advect_deposit": [c_void_p, c_double],
    "hdns_tracer_schedule": [c_void_p, c_int32, POINTER(c_int32)],
    "hdns_deposit_tracers": [c_void_p],
    "hdns_tracer_samples": [c_void_p, _D, _D],
    "hdns_tracer_trail": [c_void_p, POINTER(c_int32), POINTER(c_int32), _D],
    "hdns_tracer_trail_sample": [c_void_p, c_int32, _D],
    "hdns_tracer_trail_restore": [c_void_p, c_int32, _D, _D, c_double],
    "hdns_band_source_upload": [c_void_p, c_int32, _D],
    "hdns_below_scale": [c_void_p, _D, _D, _D],
    "hdns_tracer_time_derivative": [c_void_p, c_int32, _D],
    "hdns_tracer_cells": [c_void_p, POINTER(c_int64)],
    "hdns_traversal_order": [c_void_p, POINTER(c_int32)],
    "hdns_frame_tables": [c_void_p, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32), POINTER(c_uint8)],
    "hdns_scheme_coefficients": [c_int32, c_int32, _D, _D, _D, POINTER(c_int32)],
    "hdns_heun_weights": [c_double, _D],
    "hdns_autotune": [c_void_p, c_int32, _D, POINTER(c_int32)],
    "hdns_schedule": [c_void_p, POINTER(c_int32), POINTER(c_int32)],
    "hdns_profile_enable": [c_void_p, c_int32],
    "hdns_profile_read": [c_void_p, _D, POINTER(c_int64), c_int32],
    "hdns_synchronize": [c_void_p],
    "hdns_stream": [c_void_p, POINTER(c_void_p)],
    "hdns_step_vjp": [c_void_p, c_double, _D, _D, _D, _D],
    "hdns_adjoint_gradient": [c_void_p, c_double, c_int32, c_int32, _D, _D, _D],
    "hdns_adjoint_gradient_revolve": [c_void_p, c_double, c_int32, c_int32, _D, _D, _D],
    "hdns_revolve_schedule": [c_int32, c_int32, POINTER(c_int32), c_int64, POINTER(c_int64)],
    "hdns_ctx_create_batch": [POINTER(HdnsConfig), c_int32, _D, POINTER(c_void_p)],
    "hdns_select_box": [c_void_p, c_int32],
    "hdns_ensemble_create": [POINTER(HdnsConfig), c_int32, POINTER(c_void_p)],
    "hdns_ensemble_destroy": [c_void_p],
    "hdns_ensemble_member": [c_void_p, c_int32, POINTER(c_void_p)],
    "hdns_ensemble_step": [c_void_p, c_double],
    "hdns_ensemble_step_graph": [c_void_p, c_double],
    "hdns_ensemble_advance_tracers": [c_void_p, c_double],
    "hdns_ensemble_synchronize": [c_void_p],
}


def declared_symbols() -> list[str]:
    """Every function include/hdns.h declares."""
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:hdns_status|const char \*)\s*(hdns_\w+)\s*\(", text, flags=re.M)))


def lib() -> ctypes.CDLL:
    """The loaded library (loaded once); raises when libhdns.so has not been built."""
    global _lib
    if _lib is None:
        if not LIBPATH.exists():
            raise RuntimeError(f"{LIBPATH} is not built: run __graft_entry__.build() (hipcc, gfx950) first")
        L = ctypes.CDLL(str(LIBPATH))
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = c_int32
        L.hdns_last_error.argtypes = [c_void_p]
        L.hdns_last_error.restype = c_char_p
        L.hdns_ensemble_last_error.argtypes = [c_void_p]
        L.hdns_ensemble_last_error.restype = c_char_p
        _lib = L
    return _lib


def check(status: int, ctx=None, ensemble=None) -> None:
    """Map a status to the reference's exception classes, with the library's message."""
    if status == OK:
        return
    if ensemble is not None:
        msg = (lib().hdns_ensemble_last_error(ensemble) or b"").decode(errors="replace")
    else:
        msg = (lib().hdns_last_error(ctx) or b"").decode(errors="replace")
    if status == ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    if status == ERR_NONFINITE:
        raise BlowUpError(msg)
    if status == ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise HdnsError(msg)


def dptr(a: np.ndarray):
    """A double* to a C-contiguous float64 (or complex128, interleaved) array."""
    assert a.flags["C_CONTIGUOUS"] and a.dtype in (np.float64, np.complex128)
    return a.ctypes.data_as(_D)


def scheme_coefficients(scheme: int, precision: int):
    a = np.zeros(5)
    b = np.zeros(5)
    d = np.zeros(5)
    s = c_int32(0)
    check(lib().hdns_scheme_coefficients(scheme, precision, dptr(a), dptr(b), dptr(d), ctypes.byref(s)))
    n = s.value
    return tuple(a[:n]), tuple(b[:n]), tuple(d[:n])


def heun_weights(h: float):
    out = np.zeros(3)
    check(lib().hdns_heun_weights(h, dptr(out)))
    return tuple(out)
