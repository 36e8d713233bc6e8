"""Writes the golden fixtures of the step from the FP64 C oracle.

The reference ships no vectors for this path, so these fixtures are the FP64
oracle's outputs on the seeded workloads, frozen so that GPU tests (which must
not spend minutes re-running the FP64 oracle) and future changes to the
oracle are checked against the same numbers.  Each fixture stores a digest of
its inputs; the inputs themselves are regenerated from the seed.

    python tests/golden/make_golden.py ws1k      # WS N=1,000: 500 steps
    python tests/golden/make_golden.py ws100k    # WS N=100,000: 10,000 steps (minutes)
"""
import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from leviathan_v1_amd import workloads  # noqa: E402

DT = 0.05


def digest(wl) -> str:
    h = hashlib.sha256()
    for a in (wl.graph.row_ptr, wl.graph.col_idx, wl.delays, wl.weights, wl.theta, wl.theta_hat, wl.omega):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def make_ws1k():
    wl = workloads.small_plain(0)
    o = oracle.COracle(*wl.engine_args(), dtype=np.float64)
    out = {"digest": np.array(digest(wl)), "dt": np.array(DT)}
    r = []
    for t in range(1, 501):
        r.append(o.step(DT))
        if t in (1, 10):
            out[f"theta_{t}"] = o.theta.astype(np.float32)
            out[f"w_{t}"] = o.w.astype(np.float32)
            out[f"g_{t}"] = np.array(o.g)
    out["r"] = np.array(r)
    out["g_500"] = np.array(o.g)
    out["theta_500"] = o.theta.astype(np.float32)
    np.savez_compressed(os.path.join(HERE, "ws1k_fp64.npz"), **out)


def make_ws100k(steps=10_000):
    wl = workloads.mid_plain(0)
    o = oracle.COracle(*wl.engine_args(), dtype=np.float64)
    t0 = time.time()
    r = np.empty(steps)
    g = np.empty(steps)
    chunk = 500
    for s in range(0, steps, chunk):
        r[s:s + chunk] = o.run(DT, chunk)
        g[s + chunk - 1] = o.g
        print(f"ws100k: {s + chunk} steps, {time.time() - t0:.0f} s, r={r[s + chunk - 1]:.4f}", flush=True)
    np.savez_compressed(os.path.join(HERE, "ws100k_fp64.npz"), digest=np.array(digest(wl)), dt=np.array(DT),
                        r=r, g_final=np.array(o.g))


if __name__ == "__main__":
    oracle.build()
    for name in sys.argv[1:] or ["ws1k"]:
        {"ws1k": make_ws1k, "ws100k": make_ws100k}[name]()
