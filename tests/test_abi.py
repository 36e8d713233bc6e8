"""The C-ABI library: it loads, exports every symbol include/leviathan.h and
include/leviathan_phi.h declare,
and rejects invalid inputs with a message before touching the GPU (no compute
call is made here, so these run without one)."""
import ctypes
import os
import re

import numpy as np
import pytest

from leviathan_v1_amd import _build, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("leviathan.h", "leviathan_phi.h")]


def declared_functions():
    names = set()
    for path in HEADERS:
        text = open(path).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b((?:leviathan|phi)_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_built():
    assert os.path.exists(_build.HIP_SO), "build() must produce libleviathan_hip.so"


def test_exports_every_declared_symbol():
    lib = ctypes.CDLL(_build.HIP_SO)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each one
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_reference_entry_points_present():
    """The reference's host API (leviathan_kernel.cu:270-497) is exported under the same names."""
    lib = ctypes.CDLL(_build.HIP_SO)
    for n in ["leviathan_init", "leviathan_step", "leviathan_get_theta", "leviathan_set_theta", "leviathan_free",
              "leviathan_set_gain", "leviathan_get_gain", "leviathan_set_gain_controller",
              "leviathan_get_vram_usage", "leviathan_reset_weights"]:
        assert hasattr(lib, n)


def test_reference_phi_entry_points_present():
    """The reference's Phi host API (leviathan_phi_kernel.cu:242-368) under the same names."""
    lib = ctypes.CDLL(_build.HIP_SO)
    for n in ["phi_init", "phi_accumulate", "phi_compute", "phi_reset", "phi_get_transition_count", "phi_free"]:
        assert hasattr(lib, n)


@pytest.mark.parametrize("hubs, msg", [([], "num_hubs"), (list(range(11)), "num_hubs"), ([0, -3], "negative")])
def test_phi_init_rejects_invalid_hubs(hubs, msg):
    lib = _lib.lib()
    a = np.asarray(hubs if hubs else [0], dtype=np.int32)
    assert not lib.phi_init(len(hubs), ctypes.c_void_p(a.ctypes.data))
    assert msg in _lib.last_error()


def test_phi_null_handle_calls_report_error():
    lib = _lib.lib()
    assert np.isnan(lib.phi_compute(None))
    assert "null" in _lib.last_error()
    assert lib.phi_get_transition_count(None) == -1
    assert lib.leviathan_attach_phi(None, None) == -1


def test_default_config_matches_reference_parameters():
    cfg = _lib.default_config()
    # leviathan_kernel.cu:281-297
    expect = dict(eta=0.01, lam=0.001, gamma=0.05, alpha_eps=0.1, kappa_w=0.01, w_min=0.0, w_max=2.0, g0=1.5,
                  r0=0.5, beta=0.01, mu=0.01, g_rest=1.5, g_min=0.1, g_max=5.0)
    for k, v in expect.items():
        assert getattr(cfg, k) == np.float32(v), k
    assert cfg.gain_controller == 1


def _good():
    N = 4
    row_ptr = np.array([0, 1, 2, 3, 4], np.int32)
    col = np.array([1, 2, 3, 0], np.int32)
    dl = np.array([1, 2, 3, 4], np.uint8)
    w = np.full(4, 0.5, np.float32)
    th = np.zeros(4, np.float32)
    return [N, 5, row_ptr, col, dl, w, th, th.copy(), th.copy()]


@pytest.mark.parametrize("mutate, msg", [
    (lambda a: a[4].__setitem__(0, 9), "exceeds max_delay"),
    (lambda a: a[3].__setitem__(2, 7), "out of range"),
    (lambda a: a[2].__setitem__(2, 0), "decreases"),
    (lambda a: a.__setitem__(1, 300), "max_delay"),
])
def test_init_rejects_invalid_inputs(mutate, msg):
    lib = _lib.lib()
    a = _good()
    mutate(a)
    ptr = lambda x: ctypes.c_void_p(x.ctypes.data)
    h = lib.leviathan_init(a[0], a[1], *[ptr(x) for x in a[2:]])
    assert not h
    assert msg in _lib.last_error()


def test_engine_rejects_wrong_lengths():
    from leviathan_v1_amd import LeviathanEngine

    a = _good()
    a[6] = np.zeros(3, np.float32)
    with pytest.raises(ValueError):
        LeviathanEngine(*a)


def test_null_handle_calls_report_error():
    lib = _lib.lib()
    assert np.isnan(lib.leviathan_step(None, 0.05))
    assert "null" in _lib.last_error()
    assert lib.leviathan_step_n(None, 0.05, 3, None) == -1


def test_init_rejects_inverted_weight_bounds():
    """The step clamps every weight into [w_min, w_max]; a configuration with w_min > w_max is
    refused when the handle is created (before any device work)."""
    from leviathan_v1_amd import LeviathanEngine

    with pytest.raises(RuntimeError, match="w_min"):
        LeviathanEngine(*_good(), params=dict(w_min=1.0, w_max=0.5))
