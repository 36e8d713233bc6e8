"""``GPUPhiWorker``: hub transition statistics and the bipartition measure phi.

Mirrors the reference's pybind11 class (src/bindings/leviathan_phi_pybind.cpp:20-83):
``GPUPhiWorker(hub_indices)`` with ``accumulate(theta)``, ``compute()``, ``reset()``
and ``get_transition_count()``, the GIL released in every call (ctypes does this).

The reference feeds it from the host: every step the caller copies all N phases
off the GPU (``get_phase_snapshot``) and ``accumulate`` copies them back
(leviathan_phi_pybind.cpp:35-56).  Here ``attach(engine)`` makes every step of
the engine record the hub state of its new phases on the device, in the kernel
that finishes the step; ``accumulate`` stays for phases held on the host and
uploads only the hub phases.

Extensions: ``attach`` / ``detach``, ``accumulate_device`` (a device pointer),
``counts``, ``state``, ``sbn_tpm``, ``compute_ex`` and ``checkpoint`` / ``restore``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


class GPUPhiWorker:
    """Transition counts of the binary hub state (bit i: hub i's phase in (pi, 2pi))."""

    MAX_HUBS = 10

    def __init__(self, hub_indices):
        hubs = np.ascontiguousarray(hub_indices, dtype=np.int32).ravel()
        h = _lib.lib().phi_init(int(hubs.size), _p(hubs))
        if not h:
            raise RuntimeError("Failed to initialize GPU Phi worker: " + _lib.last_error())
        self._h = ctypes.c_void_p(h)
        self.hubs = hubs.copy()
        self.num_hubs = int(hubs.size)
        self.num_states = 1 << self.num_hubs
        self.num_partitions = (1 << (self.num_hubs - 1)) - 1
        self._engine = None

    # ------------------------------------------------------------------ reference surface
    def accumulate(self, theta) -> None:
        """Record the hub state of a host phase vector (leviathan_phi_pybind.cpp:35-56, :76-77)."""
        th = np.ascontiguousarray(theta, dtype=np.float32).ravel()
        if _lib.lib().phi_accumulate_host(self._h, _p(th), th.size) != 0:
            _lib.check("accumulate")

    def compute(self) -> float:
        """phi of the minimum-information bipartition (leviathan_phi_kernel.cu:301-338)."""
        v = _lib.lib().phi_compute(self._h)
        if v != v:
            _lib.check("compute")
        return float(v)

    def reset(self) -> None:
        _lib.lib().phi_reset(self._h)
        _lib.check("reset")

    def get_transition_count(self) -> int:
        n = _lib.lib().phi_get_transition_count(self._h)
        if n < 0:
            _lib.check("get_transition_count")
        return int(n)

    # ------------------------------------------------------------------ extensions
    def attach(self, engine) -> None:
        """Every later step of ``engine`` records the hub state of its new phases on the device."""
        if _lib.lib().leviathan_attach_phi(engine._h, self._h) != 0:
            _lib.check("attach")
        self._engine = engine

    def detach(self) -> None:
        if self._engine is not None and self._engine._h is not None:
            if _lib.lib().leviathan_attach_phi(self._engine._h, None) != 0:
                _lib.check("detach")
        self._engine = None

    def accumulate_device(self, d_theta: int) -> None:
        """Record the hub state of a phase vector in device memory (raw pointer, e.g. a
        torch tensor's ``data_ptr()``); asynchronous like the reference's phi_accumulate."""
        _lib.lib().phi_accumulate(self._h, ctypes.c_void_p(int(d_theta)))
        _lib.check("accumulate_device")

    def counts(self):
        """(counts [S, S] int64, row totals [S] int64)."""
        c = np.empty((self.num_states, self.num_states), dtype=np.int64)
        t = np.empty(self.num_states, dtype=np.int64)
        if _lib.lib().phi_get_counts(self._h, _p(c), _p(t)) != 0:
            _lib.check("counts")
        return c, t

    def state(self):
        """(last recorded state or -1, states recorded)."""
        s = ctypes.c_int()
        n = ctypes.c_int64()
        if _lib.lib().phi_get_state(self._h, ctypes.byref(s), ctypes.byref(n)) != 0:
            _lib.check("state")
        return int(s.value), int(n.value)

    def compute_ex(self) -> dict:
        """FP64 phi, the MIP mask (hub n in part A when bit n is set), the state-by-node TPM
        [S, H] and the per-bipartition values (bipartition p has mask p + 1)."""
        phi = ctypes.c_double()
        mask = ctypes.c_int()
        sbn = np.empty((self.num_states, self.num_hubs), dtype=np.float64)
        part = np.empty(max(self.num_partitions, 0), dtype=np.float64)
        if _lib.lib().phi_compute_ex(self._h, ctypes.byref(phi), ctypes.byref(mask), _p(sbn),
                                     _p(part) if part.size else None) != 0:
            _lib.check("compute_ex")
        return {"phi": float(phi.value), "mip_mask": int(mask.value), "sbn_tpm": sbn, "partitions": part}

    def sbn_tpm(self) -> np.ndarray:
        """P(hub n on at t+1 | state s at t), [S, H]; 0.5 for a state never left."""
        return self.compute_ex()["sbn_tpm"]

    def checkpoint(self) -> dict:
        c, t = self.counts()
        s, n = self.state()
        return {"hubs": self.hubs.copy(), "counts": c, "row_totals": t, "state": s, "calls": n}

    def restore(self, ck: dict) -> None:
        if not np.array_equal(np.asarray(ck["hubs"], dtype=np.int32), self.hubs):
            raise ValueError("checkpoint was taken with other hubs")
        c = np.ascontiguousarray(ck["counts"], dtype=np.int64)
        t = np.ascontiguousarray(ck["row_totals"], dtype=np.int64)
        if c.shape != (self.num_states, self.num_states) or t.shape != (self.num_states,):
            raise ValueError("checkpoint tables have the wrong shape")
        if _lib.lib().phi_set_state(self._h, _p(c), _p(t), int(ck["state"]), int(ck["calls"])) != 0:
            _lib.check("restore")

    def close(self) -> None:
        if self._h is not None:
            _lib.lib().phi_free(self._h)
            self._h = None
            self._engine = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def top_degree_hubs(row_ptr, col_idx, k: int) -> np.ndarray:
    """The k nodes of highest total degree (in + out), ties to the lower index: the hubs
    python/leviathan_iit_integration.py:217-222 picks from the networkx degree."""
    row_ptr = np.asarray(row_ptr, dtype=np.int64)
    n = row_ptr.size - 1
    deg = np.diff(row_ptr) + np.bincount(np.asarray(col_idx, dtype=np.int64), minlength=n)
    order = np.lexsort((np.arange(n), -deg))
    return order[:k].astype(np.int32)
