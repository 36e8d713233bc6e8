"""Reservoir readout benchmark.

Reservoir: Watts-Strogatz N=50,000 (k=20, p=0.1, delays uniform in [1, 50]),
omega_i(t) = omega_i + b_i u(t) with b_i ~ U(-1, 1) and u = mean of 5
sinusoids with seeded periods in [20, 200] steps plus N(0, 0.05^2) noise.
50,000 driven steps; features cos/sin of 4,096 seeded nodes (8,192 FP32
features per step).  Ridge readout (with intercept, alpha = 1e-4) predicts
u(t + 10); the first 80 % of the rows after a 1,000-step washout train it, the
rest test it.  The readout's solution is checked against an FP64 Gram/solve of
the same features on the CPU by the GPU tests (tests/test_gpu_reservoir.py);
this module only measures.
"""
from __future__ import annotations

import sys
import time

import numpy as np

from . import workloads
from .reservoir import LeviathanReservoir, multisine_input, nrmse

BF16_PEAK_TFLOPS = 2500.0


def _log(*a):
    print("[reservoir]", *a, file=sys.stderr, flush=True)


def run(N=50_000, steps=50_000, n_nodes=4096, lag=10, alpha=1e-4, washout=1000, seed=0, static=False,
        graph_chunk=100, sweep_alphas=(), sweep_washouts=()):
    """static=True freezes the reservoir: no plasticity (eta = lam = 0) and no gain controller, so the
    readout is trained and tested on a stationary system.  sweep_alphas / sweep_washouts refit the
    readout on the same captured features with other ridge parameters / washout lengths and report
    their train and test NRMSE beside the configured one."""
    t0 = time.perf_counter()
    wl = workloads.watts_strogatz_workload(N, 20, 0.1, 50, seed=seed)
    rng = np.random.Generator(np.random.PCG64(seed + 7))
    nodes = np.sort(rng.choice(N, n_nodes, replace=False)).astype(np.int32)
    b = rng.uniform(-1.0, 1.0, N).astype(np.float32)
    u = multisine_input(steps, seed=seed + 11)
    params = dict(eta=0.0, lam=0.0, gain_controller=False) if static else None
    res = LeviathanReservoir(wl, nodes, b, max_rows=steps, params=params, graph_chunk=graph_chunk)
    _log(f"setup {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    r_run = res.run(u)
    run_s = time.perf_counter() - t0
    _log(f"{steps} driven steps with capture: {run_s:.2f} s")

    rows = steps - washout - lag
    n_train = int(0.8 * rows)
    row0 = washout
    y = u[row0 + lag: row0 + lag + rows].astype(np.float64)
    fit = res.fit(y[:n_train], row0, alpha=alpha)
    _log(f"fit: {fit.iterations} refinement iterations, rel residual {fit.rel_residual:.2e}, timings {fit.timings}")
    pred = res.predict(fit, row0, rows)
    out = {
        "static_reservoir": static,
        "rows_train": n_train, "rows_test": rows - n_train,
        "nrmse_train": nrmse(pred[:n_train], y[:n_train]),
        "nrmse_test": nrmse(pred[n_train:], y[n_train:]),
        "refine_iterations": fit.iterations, "rel_residual": fit.rel_residual, "preconditioner_shift": fit.shift,
        "timings": fit.timings, "run_s": run_s,
        # context: predicting the training mean gives NRMSE ~1 on the test rows
        "mean_predictor_nrmse_test": nrmse(np.full(rows - n_train, y[:n_train].mean()), y[n_train:]),
        "r_mean_train_rows": float(np.mean(r_run[row0:row0 + n_train])),
        "r_mean_test_rows": float(np.mean(r_run[row0 + n_train:row0 + rows])),
    }

    def refit(a_, w_):
        rows_ = steps - w_ - lag
        nt = int(0.8 * rows_)
        y_ = u[w_ + lag: w_ + lag + rows_].astype(np.float64)
        f_ = res.fit(y_[:nt], w_, alpha=a_)
        p_ = res.predict(f_, w_, rows_)
        return {"alpha": a_, "washout": w_, "nrmse_train": nrmse(p_[:nt], y_[:nt]),
                "nrmse_test": nrmse(p_[nt:], y_[nt:]), "refine_iterations": f_.iterations}

    if sweep_alphas:
        out["alpha_sweep"] = [refit(float(a_), washout) for a_ in sweep_alphas]
        _log("alpha sweep", out["alpha_sweep"])
    if sweep_washouts:
        out["washout_sweep"] = [refit(alpha, int(w_)) for w_ in sweep_washouts]
        _log("washout sweep", out["washout_sweep"])
    F = 2 * n_nodes
    Fp = (F + 255) // 256 * 256
    Tp = (n_train + 63) // 64 * 64
    nI, nJ = Fp // 128, Fp // 256
    tiles = sum(nJ - i // 2 for i in range(nI))
    mfma_flops = 2.0 * Tp * 128 * 256 * tiles  # issued by the tiles computed
    # the BF16 matrix-core Gram (the first preconditioner every fit forms)
    ms = fit.timings["gram_ms_by_terms"][1]
    out["gram"] = {"flop_full_product": 2.0 * n_train * F * F, "flop_issued": mfma_flops, "kernel_ms": ms,
                   "tflops_issued": mfma_flops / (ms * 1e-3) / 1e12,
                   "roofline": {"bound": "mfma", "achieved": mfma_flops / (ms * 1e-3) / 1e12,
                                "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                                "frac": mfma_flops / (ms * 1e-3) / 1e12 / BF16_PEAK_TFLOPS}}
    if 0 in fit.timings["gram_ms_by_terms"]:
        # the FP64 Gram (rocBLAS DSYRK, lower triangle: n (n + 1) k FLOP) when the refinement needed it
        ms64 = fit.timings["gram_ms_by_terms"][0]
        fl64 = float(F) * (F + 1) * n_train
        out["gram_fp64"] = {"flop": fl64, "kernel_ms": ms64, "tflops": fl64 / (ms64 * 1e-3) / 1e12,
                            "peak_tflops": 78.6}
    res.close()
    return out
