"""Throughput of the batched ensemble (r(K) sweep) against sequential plain engines.

Workload: 256 Watts-Strogatz networks (k=20, p=0.1), sizes log-uniform in
[2,000, 200,000], delays uniform in [1, 50], gain K on linspace(0, 4, 256),
controller off.  The batched form runs every network in one fused launch per
step (uint16 history, autotuned schedule, 100-step graph chunks); the
baseline steps 256 separate plain engines one after another through the
reference-style per-step call (one synchronisation per network per step).
"""
from __future__ import annotations

import sys
import time

import numpy as np

from .engine import LeviathanEngine
from .ensemble import LeviathanEnsemble, build_workloads, sweep_spec


def _log(*a):
    print("[ensemble]", *a, file=sys.stderr, flush=True)


def run(steps: int = 200, warmup: int = 20, count: int = 256, seq_steps: int = 10, history: str = "u16",
        layout: str = "slot"):
    t0 = time.perf_counter()
    spec = sweep_spec(count)
    wls = build_workloads(spec)
    N = sum(w.N for w in wls)
    E = sum(w.E for w in wls)
    _log(f"{count} networks, N={N} E={E}, built in {time.perf_counter() - t0:.1f} s")

    ens = LeviathanEnsemble(wls, spec.gains, history=history, layout=layout, schedule="auto", graph_chunk=100)
    res = ens.engine.autotune_all(ens.dt, steps=20)
    tuned = res.get("plans_at_k", res["plans"])
    chosen = ens.engine.schedule_name
    scatter_steps = ens.engine.scatter_steps if not chosen.startswith("direct") else None
    _log("autotune", res, "->", chosen, scatter_steps)
    ens.step_n(warmup)
    t0 = time.perf_counter()
    r = ens.step_n(steps)
    el = time.perf_counter() - t0
    prof = ens.engine.profile_steps(ens.dt, 10)
    ens.close()
    batched = E * steps / el
    win = min(1000, max(1, steps // 2))
    r_mean = r[-win:].mean(axis=0)
    r_std = r[-win:].std(axis=0)
    _log(f"batched: {1e3 * el / steps:.3f} ms/step, {batched:.3e} edge-steps/s")

    # baseline: one plain engine per network, stepped one after another
    engines = [LeviathanEngine(*w.engine_args(), params=dict(g0=float(k), gain_controller=False))
               for w, k in zip(wls, spec.gains)]
    for e in engines:
        e.step(wls[0].dt)
    t0 = time.perf_counter()
    for _ in range(seq_steps):
        for e in engines:
            e.step(wls[0].dt)
    el_seq = time.perf_counter() - t0
    for e in engines:
        e.close()
    sequential = E * seq_steps / el_seq
    _log(f"sequential: {1e3 * el_seq / seq_steps:.3f} ms per ensemble step, {sequential:.3e} edge-steps/s")
    return {
        "metric": "edge-steps/s, batched ensemble of 256 networks vs 256 sequential plain engines",
        "value": batched,
        "unit": "edge-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": 1e3 * el / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded graphs and initial states)",
        "config": {"workload": f"{count} Watts-Strogatz networks k=20 p=0.1, sizes log-uniform [2000, 200000], "
                               f"K=linspace(0,4), controller off", "N": N, "E": E, "history": history, "layout": layout,
                   "schedule": chosen, "scatter_steps": scatter_steps, "graph_chunk": 100},
        "autotune_ms_per_step": tuned,
        "step_kernel_ms": prof["step_kernel_ms"] + prof["scatter_kernel_ms"] + prof["regroup_kernel_ms"],
        "sequential_plain": {"value": sequential, "unit": "edge-steps/s", "steps": seq_steps,
                             "ms_per_ensemble_step": 1e3 * el_seq / seq_steps},
        "speedup_vs_sequential": batched / sequential,
        "r_of_K": {"K": [round(float(k), 4) for k in spec.gains],
                   "r_mean": [round(float(x), 5) for x in r_mean],
                   "r_std": [round(float(x), 5) for x in r_std],
                   "window_steps": win},
    }
