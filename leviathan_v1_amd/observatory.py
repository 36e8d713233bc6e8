"""``LeviathanObservatory``: graph construction, stepping and telemetry.

Mirrors python/leviathan_h100.py:19-219.  Differences, all deliberate:

* the Watts-Strogatz graph comes from the seeded native builder instead of
  the unseeded networkx call (:34), so runs are reproducible;
* a prebuilt ``Workload`` (any graph and initial state) can be passed;
* ``step_n`` runs a chunk of steps with one synchronisation and records
  every r.
"""
from __future__ import annotations

import time
from typing import Optional

import numpy as np

from .engine import LeviathanEngine
from .workloads import Workload, watts_strogatz_workload


class LeviathanObservatory:
    """GPU-resident delayed Kuramoto network with adaptive coupling, plus telemetry."""

    def __init__(self, N: int = 100000, k: int = 20, max_delay: int = 50, seed: int = 42, *, p: float = 0.2,
                 workload: Optional[Workload] = None, reference_defaults: bool = True, verbose: bool = False,
                 **engine_kwargs):
        self.verbose = verbose
        if workload is None:
            # leviathan_h100.py:31-55: WS(N, k, p=0.2); delays randint(1, max_delay);
            # weights U(0.01, 0.1); theta U[0, 2pi); theta_hat = theta; omega ~ N(1, 0.1)
            workload = watts_strogatz_workload(N, k, p, max_delay, seed, reference_defaults=reference_defaults)
        self.workload = workload
        self.N = workload.N
        self.k = k
        self.max_delay = workload.max_delay
        self.dt = workload.dt
        self.row_ptr = workload.graph.row_ptr
        self.col_idx = workload.graph.col_idx
        self.nnz = workload.E
        self.delays = workload.delays
        self.weights = workload.weights
        self.theta = workload.theta
        self.theta_hat = workload.theta_hat
        self.omega = workload.omega
        self._log(f"[Observatory] Topology: N={self.N}, {self.nnz} edges")
        self.engine = LeviathanEngine(*workload.engine_args(), **engine_kwargs)
        self._r_capacity = 100000
        self._r_buffer = np.empty(self._r_capacity, dtype=np.float32)
        self.step_count = 0

    def _log(self, msg):
        if self.verbose:
            print(msg)

    @property
    def r_history(self):
        """Recorded r values (leviathan_h100.py:77-80)."""
        return self._r_buffer[: self.step_count]

    def _record(self, r_values):
        r_values = np.atleast_1d(np.asarray(r_values, dtype=np.float32))
        need = self.step_count + r_values.shape[0]
        if need > self._r_capacity:
            while self._r_capacity < need:
                self._r_capacity *= 2
            nb = np.empty(self._r_capacity, dtype=np.float32)
            nb[: self.step_count] = self._r_buffer[: self.step_count]
            self._r_buffer = nb
        self._r_buffer[self.step_count:need] = r_values
        self.step_count = need

    def step(self):
        """One integration step; returns r (leviathan_h100.py:82-93)."""
        r = self.engine.step(self.dt)
        self._record(r)
        return r

    def step_n(self, n: int):
        """n steps with one synchronisation; returns and records r[n]."""
        r = self.engine.step_n(self.dt, n)
        self._record(r)
        return r

    def set_gain(self, g):
        self.engine.set_gain(g)

    def get_gain(self):
        return self.engine.get_gain()

    def set_gain_controller(self, enabled):
        self.engine.set_gain_controller(enabled)

    def get_vram_usage(self):
        return self.engine.get_vram_usage()

    def reset_weights(self, weights=None):
        if weights is None:
            weights = self.weights
        self.engine.reset_weights(weights)

    def get_vram_report(self):
        """Formatted device-memory report (leviathan_h100.py:112-138)."""
        used = self.get_vram_usage()
        hist_bytes = 2 if self.engine.history == "u16" else 4
        slots = self.max_delay + 2
        breakdown = {
            "History Buffer": self.N * slots * hist_bytes,
            "Topology (CSR)": (self.N + 1) * 4 + self.nnz * (4 + 1 + 4),
            "Node States": self.N * 4 * 4,
        }
        out = f"VRAM Usage Report (N={self.N}, E={self.nnz})\n" + "-" * 40 + "\n"
        for k, v in breakdown.items():
            out += f"{k:20s} : {v / (1024 * 1024):8.2f} MB\n"
        out += "-" * 40 + "\n" + f"{'TOTAL (Measured)':20s} : {used / (1024 * 1024):8.2f} MB\n"
        return out

    def run_baseline(self, num_steps=500, log_interval=100):
        """Warm-up run (leviathan_h100.py:140-155)."""
        start = time.time()
        for i in range(num_steps):
            r = self.step()
            if self.verbose and (i + 1) % log_interval == 0:
                print(f"  Step {i + 1:5d}: r={r:.4f} ({(i + 1) / (time.time() - start):.1f} FPS)")
        return self.r_history

    def run_with_stimulus(self, num_steps=200, stimulus_start=100, stimulus_nodes=None, stimulus_strength=2.0):
        """Phase-kick protocol (leviathan_h100.py:157-196)."""
        if stimulus_nodes is None:
            stimulus_nodes = list(range(10))
        for i in range(num_steps):
            self.step()
            if i == stimulus_start:
                theta = self.get_phase_snapshot()
                theta[np.asarray(stimulus_nodes)] += stimulus_strength
                theta = theta % (2 * np.pi)
                self.set_phase_snapshot(theta)
        return self.r_history[-num_steps:]

    def get_phase_snapshot(self):
        """Phases from device memory (leviathan_h100.py:198-200)."""
        return self.engine.get_theta()

    def set_phase_snapshot(self, theta):
        """Upload phases; the delay history is reset (leviathan_h100.py:202-207)."""
        theta = np.asarray(theta)
        if theta.dtype != np.float32:
            theta = theta.astype(np.float32)
        self.engine.set_theta(theta)

    def statistics(self):
        """Telemetry statistics (leviathan_h100.py:209-219)."""
        r_arr = np.array(self.r_history)
        return {
            "final_r": r_arr[-1],
            "mean_r": np.mean(r_arr),
            "max_r": np.max(r_arr),
            "min_r": np.min(r_arr),
            "std_r": np.std(r_arr),
            "steps": len(r_arr),
        }

    def close(self):
        self.engine.close()
