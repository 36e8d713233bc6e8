"""Batched ensembles: many independent networks of uneven size in one launch.

Counterpart of the reference's gain sweep (experiments/phase_sweep.py:25-59),
which steps ONE network through 30 gains sequentially with a host
synchronisation per step.  Here every network of the ensemble has its own
gain K (controller off) and all of them advance together: one fused step
launch (plus one FP64 finalize) per step for the whole ensemble, in
HIP-graph-captured chunks, with the schedule chosen by the autotuner.  The
result is the synchronisation curve r(K).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import graph as _graph
from . import workloads as _workloads
from .engine import LeviathanEngine


@dataclass
class EnsembleSpec:
    sizes: np.ndarray        # nodes per network
    gains: np.ndarray        # K per network
    k: int = 20
    p: float = 0.1
    max_delay: int = 50
    seed: int = 0


def sweep_spec(count: int = 256, lo: int = 2000, hi: int = 200_000, k_max: float = 4.0, seed: int = 0,
               k: int = 20, p: float = 0.1, max_delay: int = 50) -> EnsembleSpec:
    """Sizes log-uniform in [lo, hi]; gains on a linear sweep 0 .. k_max."""
    return EnsembleSpec(_workloads.ensemble_sizes(count, lo, hi, seed), np.linspace(0.0, k_max, count), k, p,
                        max_delay, seed)


def build_workloads(spec: EnsembleSpec):
    return [_workloads.watts_strogatz_workload(int(n), spec.k, spec.p, spec.max_delay, seed=spec.seed * 100003 + i)
            for i, n in enumerate(spec.sizes)]


class LeviathanEnsemble:
    """Independent networks stepped together; ``r`` is per network."""

    def __init__(self, wls, gains: Sequence[float], controller: Optional[Sequence[bool]] = None, **engine_kwargs):
        if len(wls) != len(gains):
            raise ValueError("one gain per network")
        md = {w.max_delay for w in wls}
        if len(md) != 1:
            raise ValueError("all networks must share max_delay")
        dt = {w.dt for w in wls}
        if len(dt) != 1:
            raise ValueError("all networks must share dt")
        self.dt = dt.pop()
        g, offs = _graph.disjoint_union([w.graph for w in wls])
        cat = lambda f: np.concatenate([getattr(w, f) for w in wls])
        self.offsets = offs
        self.n_nets = len(wls)
        self.N = g.num_nodes
        self.E = g.num_edges
        ctl = [False] * self.n_nets if controller is None else list(controller)
        self.engine = LeviathanEngine(self.N, md.pop(), g.row_ptr, g.col_idx, cat("delays"), cat("weights"),
                                      cat("theta"), cat("theta_hat"), cat("omega"), n_nets=self.n_nets,
                                      net_offsets=offs, g0=list(gains), controller=ctl, **engine_kwargs)
        self.gains = np.asarray(gains, dtype=np.float64)

    def step_n(self, n: int) -> np.ndarray:
        """n steps; returns r as float32[n, n_nets]."""
        return self.engine.step_n(self.dt, n)

    def phase_sweep(self, warmup: int, measure: int):
        """r(K): mean and std of r per network over `measure` steps after `warmup` steps
        (phase_sweep.py:42-55 measures 100 steps after 200 warm-up steps per gain)."""
        if warmup:
            self.step_n(warmup)
        r = self.step_n(measure)
        return self.gains.copy(), r.mean(axis=0), r.std(axis=0)

    def theta(self, net: int) -> np.ndarray:
        th = self.engine.get_theta()
        return th[self.offsets[net]:self.offsets[net + 1]]

    def close(self):
        self.engine.close()
