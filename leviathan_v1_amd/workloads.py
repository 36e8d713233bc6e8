"""Seeded synthetic workloads of the step.

Each workload is a graph plus the initial state the engine is created from.
The distributions are the ones the measured configurations name:

* delays uniform integers in [1, max_delay] (uint8),
* weights uniform [0, 1) float32, phases uniform [0, 2*pi) float32,
* natural frequencies omega ~ N(0, 1) float32, theta_hat0 = theta0.

The reference's own defaults (python/leviathan_h100.py:51-55: delays in
[1, max_delay-1], weights U(0.01, 0.1), omega ~ N(1, 0.1), WS p=0.2) are
available as ``reference_defaults=True``.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import graph as _graph


@dataclass
class Workload:
    name: str
    graph: _graph.CSRGraph
    max_delay: int
    delays: np.ndarray
    weights: np.ndarray
    theta: np.ndarray
    theta_hat: np.ndarray
    omega: np.ndarray
    dt: float = 0.05
    meta: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return self.graph.num_nodes

    @property
    def E(self) -> int:
        return self.graph.num_edges

    def engine_args(self):
        """Positional arguments of LeviathanEngine(...) (leviathan_pybind11.cpp:32-35)."""
        return (self.N, self.max_delay, self.graph.row_ptr, self.graph.col_idx, self.delays,
                self.weights, self.theta, self.theta_hat, self.omega)


def relabel(wl: Workload, perm, name: str | None = None) -> Workload:
    """The same workload with node i of the result being node perm[i] of wl: graph, delays,
    weights and node states carried along (see graph.permute, degree_order, rcm_order)."""
    perm = np.asarray(perm, dtype=np.int32)
    g, emap = _graph.permute(wl.graph, perm)
    return Workload(name or f"{wl.name} relabelled", g, wl.max_delay, wl.delays[emap], wl.weights[emap],
                    wl.theta[perm], wl.theta_hat[perm], wl.omega[perm], wl.dt, dict(wl.meta, relabelled=True))


def init_state(g: _graph.CSRGraph, max_delay: int, seed: int, reference_defaults: bool = False):
    rng = np.random.Generator(np.random.PCG64(seed))
    E, N = g.num_edges, g.num_nodes
    if reference_defaults:
        delays = rng.integers(1, max_delay, E, dtype=np.uint8)
        weights = rng.uniform(0.01, 0.1, E).astype(np.float32)
    else:
        delays = rng.integers(1, max_delay + 1, E, dtype=np.uint8)
        weights = rng.random(E, dtype=np.float32)
    theta = rng.random(N, dtype=np.float32) * np.float32(2.0 * np.pi)
    if reference_defaults:
        omega = rng.normal(1.0, 0.1, N).astype(np.float32)
    else:
        omega = rng.standard_normal(N, dtype=np.float32)
    return delays, weights, theta, theta.copy(), omega


def watts_strogatz_workload(n: int, k: int = 20, p: float = 0.1, max_delay: int = 50, seed: int = 0,
                            name: str | None = None, reference_defaults: bool = False) -> Workload:
    g = _graph.watts_strogatz(n, k, p, seed)
    d, w, th, thh, om = init_state(g, max_delay, seed + 1, reference_defaults)
    return Workload(name or f"ws_n{n}_k{k}_p{p}", g, max_delay, d, w, th, thh, om,
                    meta={"graph": "watts_strogatz", "n": n, "k": k, "p": p, "seed": seed})


def barabasi_albert_workload(n: int, m: int = 10, max_delay: int = 50, seed: int = 0,
                             name: str | None = None) -> Workload:
    g = _graph.barabasi_albert(n, m, seed)
    d, w, th, thh, om = init_state(g, max_delay, seed + 1)
    return Workload(name or f"ba_n{n}_m{m}", g, max_delay, d, w, th, thh, om,
                    meta={"graph": "barabasi_albert", "n": n, "m": m, "seed": seed})


def small_plain(seed: int = 0) -> Workload:
    """WS N=1,000 k=20 p=0.1, delays [1,50], 500 steps at dt=0.05."""
    return watts_strogatz_workload(1000, 20, 0.1, 50, seed, name="ws1k")


def mid_plain(seed: int = 0) -> Workload:
    """WS N=100,000 k=20 p=0.1, delays [1,50], 10,000 steps."""
    return watts_strogatz_workload(100_000, 20, 0.1, 50, seed, name="ws100k")


def scale_skew(seed: int = 0, n: int = 10_000_000) -> Workload:
    """Barabasi-Albert N=10,000,000 m=10 (heavy-tailed in-degree), delays [1,50]."""
    return barabasi_albert_workload(n, 10, 50, seed, name=f"ba{n}")


def ensemble_sizes(count: int = 256, lo: int = 2000, hi: int = 200_000, seed: int = 0) -> np.ndarray:
    """Network sizes log-uniform in [lo, hi]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return np.exp(rng.uniform(np.log(lo), np.log(hi), count)).astype(np.int64)
