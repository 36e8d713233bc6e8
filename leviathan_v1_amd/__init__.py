"""leviathan_v1_amd -- LEVIATHAN's core step on one MI355X (gfx950 HIP).

Public surface (mirrors the reference's Python layer):

* ``LeviathanEngine``      -- src/bindings/leviathan_pybind11.cpp:82-104
* ``LeviathanObservatory`` -- python/leviathan_h100.py:19-219
* ``phi.GPUPhiWorker``     -- src/bindings/leviathan_phi_pybind.cpp:71-83 (hub transition statistics)

The compute lives in ``libleviathan_hip.so`` (C-ABI: include/leviathan.h).
Importing the engine classes loads it and fails loudly if it is missing.
"""
from . import graph, workloads  # noqa: F401  (host-only helpers)

__all__ = ["LeviathanEngine", "LeviathanObservatory", "graph", "workloads"]


def __getattr__(name):
    if name == "LeviathanEngine":
        from .engine import LeviathanEngine

        return LeviathanEngine
    if name == "LeviathanObservatory":
        from .observatory import LeviathanObservatory

        return LeviathanObservatory
    raise AttributeError(name)
