"""ctypes binding of ``libleviathan_hip.so`` (declarations: include/leviathan.h and
include/leviathan_phi.h).

``ctypes.CDLL`` releases the GIL for the duration of every foreign call, so
``step`` (and every other engine call) runs with the GIL released, as the
reference's pybind11 ``step`` does (src/bindings/leviathan_pybind11.cpp:44-47).

There is no fallback: if the library is missing or fails to load, importing
the engine raises.
"""
from __future__ import annotations

import ctypes
import os

from . import _build

c_int = ctypes.c_int
c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_float = ctypes.c_float
c_double = ctypes.c_double
c_bool = ctypes.c_bool
c_size_t = ctypes.c_size_t
vp = ctypes.c_void_p


class LeviathanConfig(ctypes.Structure):
    """Mirror of ``LeviathanConfig`` in include/leviathan.h."""

    _fields_ = [
        ("eta", c_float), ("lam", c_float), ("gamma", c_float), ("alpha_eps", c_float),
        ("kappa_w", c_float), ("w_min", c_float), ("w_max", c_float),
        ("g0", c_float), ("r0", c_float), ("beta", c_float), ("mu", c_float),
        ("g_rest", c_float), ("g_min", c_float), ("g_max", c_float),
        ("gain_controller", c_int32),
        ("history", c_int32), ("layout", c_int32), ("edges_per_thread", c_int32),
        ("graph_chunk", c_int32),
        ("schedule", c_int32), ("bucket_edges", c_int32),
    ]


HISTORY = {"f32": 0, "u16": 1}
LAYOUT = {"node": 0, "slot": 1, "auto": 2}
SCHEDULE = {"direct": 0, "binned": 1, "auto": 2, "staged": 3}
PLAN_NAMES = [f"{k}/{e}" for k in ("direct", "binned", "staged") for e in (4, 8, 16)]

# name -> (restype, argtypes); every symbol include/leviathan.h and include/leviathan_phi.h declare.
SIGNATURES = {
    "leviathan_init": (vp, [c_int, c_int, vp, vp, vp, vp, vp, vp, vp]),
    "leviathan_step": (c_float, [vp, c_float]),
    "leviathan_get_theta": (None, [vp, vp]),
    "leviathan_set_theta": (None, [vp, vp]),
    "leviathan_free": (None, [vp]),
    "leviathan_set_gain": (None, [vp, c_float]),
    "leviathan_get_gain": (c_float, [vp]),
    "leviathan_set_gain_controller": (None, [vp, c_bool]),
    "leviathan_get_vram_usage": (c_size_t, [vp]),
    "leviathan_reset_weights": (None, [vp, vp]),
    "leviathan_last_error": (ctypes.c_char_p, []),
    "leviathan_config_default": (None, [ctypes.POINTER(LeviathanConfig)]),
    "leviathan_init_ex": (vp, [ctypes.POINTER(LeviathanConfig), c_int, c_int, vp, vp, vp, vp, vp, vp, vp,
                               c_int, vp, vp, vp]),
    "leviathan_step_n": (c_int, [vp, c_float, c_int, vp]),
    "leviathan_set_gain_net": (c_int, [vp, c_int, c_float]),
    "leviathan_get_gain_net": (c_float, [vp, c_int]),
    "leviathan_set_gain_controller_net": (c_int, [vp, c_int, c_bool]),
    "leviathan_num_nets": (c_int, [vp]),
    "leviathan_get_weights": (c_int, [vp, vp]),
    "leviathan_get_predictor": (c_int, [vp, vp, vp]),
    "leviathan_get_history": (c_int, [vp, c_int, vp]),
    "leviathan_set_predictor": (c_int, [vp, vp, vp]),
    "leviathan_set_history": (c_int, [vp, c_int, vp]),
    "leviathan_get_gains": (c_int, [vp, vp]),
    "leviathan_set_gains": (c_int, [vp, vp]),
    "leviathan_set_step_count": (c_int, [vp, c_int64]),
    "leviathan_num_nodes": (c_int, [vp]),
    "leviathan_num_edges": (c_int64, [vp]),
    "leviathan_step_count": (c_int64, [vp]),
    "leviathan_num_blocks": (c_int, [vp]),
    "leviathan_profile_steps": (c_int, [vp, c_float, c_int, vp]),
    "leviathan_set_drive": (c_int, [vp, vp]),
    "leviathan_set_schedule": (c_int, [vp, c_int, c_int]),
    "leviathan_get_schedule": (c_int, [vp, vp, vp]),
    "leviathan_autotune": (c_int, [vp, c_float, c_int, vp]),
    "leviathan_set_scatter_steps": (c_int, [vp, c_int]),
    "leviathan_get_scatter_steps": (c_int, [vp]),
    "leviathan_autotune_scatter_steps": (c_int, [vp, c_float, c_int, vp]),
    "leviathan_set_row_steps": (c_int, [vp, c_int]),
    "leviathan_get_row_steps": (c_int, [vp]),
    "leviathan_get_layout": (c_int, [vp]),
    "leviathan_step_n_driven": (c_int, [vp, c_float, c_int, vp, vp]),
    "leviathan_features_setup": (c_int, [vp, vp, c_int, c_int]),
    "leviathan_step_n_features": (c_int, [vp, c_float, c_int, vp, vp]),
    "leviathan_features_rows": (c_int64, [vp]),
    "leviathan_features_get": (c_int, [vp, c_int64, c_int64, vp]),
    "leviathan_features_colsum": (c_int, [vp, c_int64, c_int64, vp, vp]),
    "leviathan_features_colsum_noalloc": (c_int, [vp, c_int64, c_int64, vp, vp]),
    "leviathan_features_matvec": (c_int, [vp, c_int64, c_int64, vp, vp, vp]),
    "leviathan_features_gram": (c_int, [vp, c_int64, c_int64, vp, c_int, vp, vp]),
    "leviathan_features_factor": (c_int, [vp, c_int64, c_int64, vp, c_int, c_double, vp, vp]),
    "leviathan_features_precond": (c_int, [vp, vp, vp]),
    # include/leviathan_phi.h
    "phi_init": (vp, [c_int, vp]),
    "phi_accumulate": (None, [vp, vp]),
    "phi_compute": (c_float, [vp]),
    "phi_reset": (None, [vp]),
    "phi_get_transition_count": (c_int64, [vp]),
    "phi_free": (None, [vp]),
    "leviathan_attach_phi": (c_int, [vp, vp]),
    "phi_accumulate_host": (c_int, [vp, vp, c_int64]),
    "phi_get_counts": (c_int, [vp, vp, vp]),
    "phi_get_state": (c_int, [vp, vp, vp]),
    "phi_set_state": (c_int, [vp, vp, vp, c_int, c_int64]),
    "phi_compute_ex": (c_int, [vp, vp, vp, vp, vp]),
}

_lib = None


def lib():
    """Load libleviathan_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        path = _build.HIP_SO
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is missing: the gfx950 library must be built first "
                "(python -m leviathan_v1_amd._build hip)"
            )
        handle = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def last_error() -> str:
    return lib().leviathan_last_error().decode(errors="replace")


def check(what: str = "") -> None:
    msg = last_error()
    if msg:
        raise RuntimeError(f"{what}: {msg}" if what else msg)


def default_config() -> LeviathanConfig:
    cfg = LeviathanConfig()
    lib().leviathan_config_default(ctypes.byref(cfg))
    return cfg
