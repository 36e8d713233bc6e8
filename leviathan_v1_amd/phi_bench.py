"""Cost of the hub transition statistics on the 10^7-node Barabasi-Albert step.

Workload: Barabasi-Albert N = 10^7, m = 10, delays uniform in [1, 50], uint16
history, staged schedule, 100-step graph chunks; the 10 hubs of highest degree.

* attached: ms per step with a PhiHandle attached (every step records the hub
  state on the device, in the finalize kernel) against the same engine without;
* round trip: the reference's per-step path (leviathan_iit_integration.py:237-244
  with leviathan_phi_pybind.cpp:35-56): copy the N phases to the host, upload
  them again and run the accumulation, timed per step on top of the step;
* compute: phi_compute over the S = 1024-state table (wall time per call).

Prints one JSON line.  Run: python -m leviathan_v1_amd.phi_bench
"""
from __future__ import annotations

import ctypes
import json
import sys
import time

import numpy as np

from . import workloads
from .engine import LeviathanEngine
from .phi import GPUPhiWorker, top_degree_hubs


def _log(*a):
    print("[phi]", *a, file=sys.stderr, flush=True)


def _time_steps(eng, dt, steps):
    eng.step_n(dt, 100)
    t0 = time.perf_counter()
    eng.step_n(dt, steps)
    return 1e3 * (time.perf_counter() - t0) / steps


def run(n: int = 10_000_000, steps: int = 400, trips: int = 20, computes: int = 50, hubs: int = 10):
    t0 = time.perf_counter()
    wl = workloads.barabasi_albert_workload(n, 10, 50, seed=0)
    _log(f"N={wl.N} E={wl.E} built in {time.perf_counter() - t0:.1f} s")
    dt = wl.dt
    hub_idx = top_degree_hubs(wl.graph.row_ptr, wl.graph.col_idx, hubs)
    eng = LeviathanEngine(*wl.engine_args(), history="u16", schedule="staged")
    w = GPUPhiWorker(hub_idx)

    plain = [_time_steps(eng, dt, steps)]
    w.attach(eng)
    attached = [_time_steps(eng, dt, steps)]
    w.detach()
    plain.append(_time_steps(eng, dt, steps))
    w.attach(eng)
    attached.append(_time_steps(eng, dt, steps))
    recorded = w.get_transition_count()
    _log(f"plain {plain} ms/step, attached {attached} ms/step, {recorded} states recorded")
    w.detach()

    # the reference's host round trip per step: N phases down, N phases up, accumulate
    hip = ctypes.CDLL("libamdhip64.so")
    d = ctypes.c_void_p()
    host = np.empty(wl.N, np.float32)
    assert hip.hipMalloc(ctypes.byref(d), ctypes.c_size_t(host.nbytes)) == 0
    w.reset()
    try:
        eng.step(dt)
        t0 = time.perf_counter()
        for _ in range(trips):
            eng.step(dt)
            host[:] = eng.get_theta()
            assert hip.hipMemcpy(d, ctypes.c_void_p(host.ctypes.data), ctypes.c_size_t(host.nbytes), 1) == 0
            w.accumulate_device(d.value)
        w.get_transition_count()  # waits for the last accumulation
        trip_total = 1e3 * (time.perf_counter() - t0) / trips
        t0 = time.perf_counter()
        for _ in range(trips):
            eng.step(dt)
        step_alone = 1e3 * (time.perf_counter() - t0) / trips
    finally:
        hip.hipFree(d)
    _log(f"round trip {trip_total:.3f} ms/step with the step, step alone {step_alone:.3f} ms")

    # phi_compute on a table that has seen every attached step above
    w.attach(eng)
    eng.step_n(dt, 1000)
    w.compute()
    t0 = time.perf_counter()
    for _ in range(computes):
        w.compute()
    compute_ms = 1e3 * (time.perf_counter() - t0) / computes
    ex = w.compute_ex()
    states_seen = int((w.counts()[1] > 0).sum())
    w.close()
    eng.close()
    out = {
        "workload": f"barabasi_albert N={wl.N} m=10 E={wl.E} delays 1..50, u16 history, staged schedule",
        "hubs": hub_idx.tolist(),
        "ms_per_step_plain": min(plain),
        "ms_per_step_attached": min(attached),
        "attached_overhead_ms": min(attached) - min(plain),
        "reference_round_trip_ms_per_step": trip_total - step_alone,
        "single_step_ms": step_alone,
        "phi_compute_ms": compute_ms,
        "phi": ex["phi"],
        "mip_mask": ex["mip_mask"],
        "states_seen": states_seen,
    }
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    run()
