"""Checkpoint files of a LeviathanEngine's full state.

A checkpoint (LeviathanEngine.checkpoint()) holds everything the next steps
depend on: theta, theta_hat, omega_hat, the weights, the history values seen by
delays 1..max_delay, each network's FP64 gain and the step count, plus the
graph's sizes and a digest of its CSR arrays.  Restoring it into an engine
built on the same graph continues the run exactly (the reference can only set
theta, which erases the history: leviathan_kernel.cu:445-453).

Files are NumPy .npz archives of plain arrays (no pickled objects).
"""
from __future__ import annotations

import hashlib

import numpy as np

FIELDS = ("theta", "theta_hat", "omega_hat", "weights", "history", "gains", "step", "N", "E", "max_delay")


def graph_digest(row_ptr, col_idx, delays) -> str:
    """SHA-256 of the CSR arrays and delays in the engine's own types (int32, int32, uint8),
    so the same graph digests the same whatever integer types the caller holds it in."""
    h = hashlib.sha256()
    for a, t in ((row_ptr, np.int32), (col_idx, np.int32), (delays, np.uint8)):
        a = np.asarray(a)
        if a.size and a.dtype != t and not np.array_equal(a, a.astype(t)):
            raise ValueError(f"graph array does not fit {np.dtype(t).name}")
        h.update(memoryview(np.ascontiguousarray(a, dtype=t)).cast("B"))
    return h.hexdigest()


def save(path, state: dict, row_ptr=None, col_idx=None, delays=None) -> None:
    """Write a checkpoint; with the CSR arrays and delays it also records their digest."""
    missing = [k for k in FIELDS if k not in state]
    if missing:
        raise ValueError(f"checkpoint lacks {missing}")
    extra = {}
    if row_ptr is not None:
        extra["graph_digest"] = np.array(graph_digest(row_ptr, col_idx, delays))
    elif "graph_digest" in state:
        extra["graph_digest"] = np.asarray(state["graph_digest"])
    np.savez(path, **{k: np.asarray(state[k]) for k in FIELDS}, **extra)


def load(path, row_ptr=None, col_idx=None, delays=None) -> dict:
    """Read a checkpoint; with the CSR arrays and delays it checks the recorded digest."""
    with np.load(path, allow_pickle=False) as z:
        state = {k: z[k] for k in z.files}
    if row_ptr is not None and "graph_digest" in state:
        if str(state["graph_digest"]) != graph_digest(row_ptr, col_idx, delays):
            raise ValueError("checkpoint was saved from a different graph")
    return state
