#!/usr/bin/env python3
"""Throughput of LEVIATHAN's fused step on one MI355X.

Default workload: Barabasi-Albert graph N=10,000,000, m=10 (heavy-tailed
in-degree, ~2.0e8 incoming edges), delays uniform in [1, 50], FP32 weights
and phases.  The step is run in three forms from the same seeded inputs
already resident in HBM: the uint16 exact-wrap history with the autotuned
schedule (the headline form: ``value``), the plain FP32 drop-in exactly as
``leviathan_init`` configures it (FP32 history, direct schedule, 8 edges per
lane: ``plain_fp32``), and the FP32 history with the autotuned schedule
(``fp32_tuned``).

One JSON line on stdout.  ``value`` = edge-steps/s (true nnz x steps / s) of
the uint16 form; ``hbm_gbs`` = algorithmic bytes per step x steps/s.  The
``roofline`` object describes the fused step kernel (HIP-event durations,
measured live); ``cpu_baseline`` times the FP32 OpenMP CPU port of the
reference's step (oracle/) on a bounded number of steps of the same graph.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload ba10m|ws100k|ensemble]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "edge-steps/s + HBM GB/s: uint16-delay/batched vs plain fp32, N=10M k=20"
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes of one fused step: each logically required byte once.
#   per edge: col_idx 4 + delay 1 + weight read+write 8 + delayed-phase gather (4 FP32 | 2 uint16)
#   per node: row_ptr 4 + theta r+w 8 + theta_hat r+w 8 + omega_hat r+w 8 + omega 4 + history write (4 | 2)
BYTES_PER_EDGE = {"f32": 17, "u16": 15}
BYTES_PER_NODE = {"f32": 36, "u16": 34}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def algorithmic_bytes(N, E, history):
    return BYTES_PER_EDGE[history] * E + BYTES_PER_NODE[history] * N


def make_workload(name, n_override=None):
    from leviathan_v1_amd import workloads

    if name == "ba10m":
        n = n_override or 10_000_000
        return workloads.barabasi_albert_workload(n, 10, 50, seed=0,
                                                  name=f"barabasi_albert N={n:,} m=10 delays U{{1..50}}")
    if name == "ws100k":
        n = n_override or 100_000
        return workloads.watts_strogatz_workload(n, 20, 0.1, 50, seed=0,
                                                 name=f"watts_strogatz N={n:,} k=20 p=0.1 delays U{{1..50}}")
    raise ValueError(name)


def run_form(wl, history, steps, warmup, layout="node", ept=8, schedule="auto", bucket_edges=0, profile_steps=20,
             scatter_steps="auto", row_steps=None):
    from leviathan_v1_amd import LeviathanEngine

    t0 = time.perf_counter()
    eng = LeviathanEngine(*wl.engine_args(), history=history, layout=layout, edges_per_thread=ept,
                          schedule=schedule, bucket_edges=bucket_edges)
    if row_steps:
        eng.set_row_steps(row_steps)
    log(f"{history}: engine ready in {time.perf_counter() - t0:.1f} s")
    try:
        tuned = k_ms = None
        if schedule == "auto" and scatter_steps == "auto":
            # plans, steps per scatter pass of the best two-phase plan, plans again at that K
            res = eng.autotune_all(wl.dt, steps=20)
            tuned, k_ms = res.get("plans_at_k", res["plans"]), res.get("scatter_steps")
            log(f"{history}: autotune {res} -> {eng.get_schedule()}, {eng.scatter_steps} steps per pass")
        else:
            tuned = eng.autotune(wl.dt, steps=5) if schedule == "auto" else None
            if eng.get_schedule()[0] != "direct":
                if scatter_steps == "auto":
                    k_ms = eng.autotune_scatter_steps(wl.dt, steps=8)
                else:
                    eng.set_scatter_steps(int(scatter_steps))
            log(f"{history}: autotune {tuned}, scatter steps {k_ms} -> {eng.get_schedule()}, "
                f"{eng.scatter_steps} steps per pass")
        chosen = eng.get_schedule()
        eng.step_n(wl.dt, warmup)
        t0 = time.perf_counter()
        r = eng.step_n(wl.dt, steps)  # returns after the stream is synchronised
        el = time.perf_counter() - t0
        prof = eng.profile_steps(wl.dt, profile_steps)
        theta = eng.get_theta()
        ok = bool(np.isfinite(theta).all() and np.isfinite(r).all() and 0 <= r.min() and r.max() <= 1)
        return {
            "history": history, "layout": eng.layout, "schedule": eng.schedule_name,
            "autotune_ms_per_step": tuned, "scatter_steps": eng.scatter_steps if chosen[0] != "direct" else None,
            "row_steps": eng.row_steps,
            "scatter_steps_ms_per_step": k_ms,
            "seconds": el, "ms_per_step": 1e3 * el / steps,
            "edge_steps_per_s": wl.E * steps / el,
            "hbm_gbs": algorithmic_bytes(wl.N, wl.E, history) * steps / el / 1e9,
            "step_kernel_ms": prof["step_kernel_ms"], "finalize_kernel_ms": prof["finalize_kernel_ms"],
            "scatter_kernel_ms": prof["scatter_kernel_ms"], "regroup_kernel_ms": prof["regroup_kernel_ms"],
            "r_last": float(r[-1]), "invariants_ok": ok, "vram_bytes": eng.get_vram_usage(),
            "row_blocks": eng.num_blocks,
        }
    finally:
        eng.close()


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(wl, budget_s=10.0, max_steps=50):
    """FP32 OpenMP port of the reference's four-kernel step (oracle/), same graph and inputs,
    timed with all the threads OpenMP is given and with one thread."""
    import oracle

    oracle.build()
    o = oracle.COracle(*wl.engine_args(), dtype=np.float32)
    o.run(wl.dt, 1)  # first touch

    def rate(threads):
        prev = oracle.set_threads(threads)
        try:
            o.run(wl.dt, 1)
            n, t0 = 0, time.perf_counter()
            while n < max_steps and time.perf_counter() - t0 < budget_s:
                o.run(wl.dt, 1)
                n += 1
            return wl.E * n / (time.perf_counter() - t0), n, time.perf_counter() - t0
        finally:
            oracle.set_threads(prev)

    threads = oracle.set_threads(1)
    oracle.set_threads(threads)
    all_rate, n_all, el_all = rate(threads)
    one_rate, n_one, el_one = rate(1)
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = os.cpu_count()
    return {"value": all_rate, "unit": "edge-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n_all} steps at {threads} threads ({el_all:.1f} s) and {n_one} steps at 1 thread "
                      f"({el_one:.1f} s) of the same graph and inputs (FP32 C/OpenMP port of "
                      f"leviathan_kernel.cu:93-264,391-438)",
            "value_1_thread": one_rate, "nproc": os.cpu_count(), "affinity_cpus": affinity,
            "cpu_model": _cpu_model()}


def ws100k_record(steps=10_000, warmup=100):
    """Watts-Strogatz N=100,000 (k=20, p=0.1, delays U{1..50}), plasticity and the gain
    controller on: the plain FP32 drop-in (leviathan_init's configuration) over 10,000 steps in
    100-step graph chunks, and its windowed r against the FP64 oracle's 10,000 steps of the same
    inputs (the committed fixture tests/golden/ws100k_fp64.npz, when present)."""
    from leviathan_v1_amd import LeviathanEngine

    wl = make_workload("ws100k")
    eng = LeviathanEngine(*wl.engine_args())
    try:
        eng.step_n(wl.dt, warmup)
        eng.close()
        eng = LeviathanEngine(*wl.engine_args())  # the measured run starts from the inputs
        t0 = time.perf_counter()
        r = eng.step_n(wl.dt, steps)
        el = time.perf_counter() - t0
        prof = eng.profile_steps(wl.dt, 50)
    finally:
        eng.close()
    out = {"workload": wl.name, "N": wl.N, "E": wl.E, "form": "plain FP32 drop-in (direct/8)", "steps": steps,
           "ms_per_step": 1e3 * el / steps, "edge_steps_per_s": wl.E * steps / el,
           "hbm_gbs": algorithmic_bytes(wl.N, wl.E, "f32") * steps / el / 1e9,
           "step_kernel_ms": prof["step_kernel_ms"], "finalize_kernel_ms": prof["finalize_kernel_ms"],
           "note": "the 41 MB working set fits the 256 MB Infinity Cache: algorithmic GB/s, not HBM-bound"}
    gold = os.path.join(ROOT, "tests", "golden", "ws100k_fp64.npz")
    if os.path.exists(gold) and steps == 10_000:
        ref = np.load(gold)["r"]
        win = lambda x: np.asarray(x, np.float64)[1000:].reshape(-1, 500)
        out["windowed_r_vs_fp64"] = {"window": 500, "warmup": 1000,
                                     "max_abs_mean_diff": float(np.abs(win(r).mean(1) - win(ref).mean(1)).max()),
                                     "max_abs_var_diff": float(np.abs(win(r).var(1) - win(ref).var(1)).max())}
    return out


def copy_calibration(nbytes: int, reps: int = 3) -> None:
    """A plain device-to-device copy of known size (HIP runtime blit kernel), so that
    profiler byte counters can be scaled: counter(copy) / (2 * nbytes * reps)."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(b),
                                                                                  ctypes.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    try:
        hip.hipMemset(a, 1, ctypes.c_size_t(nbytes))
        for _ in range(reps):
            if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3):  # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        hip.hipDeviceSynchronize()
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    log(f"copy calibration: {reps} x {nbytes} bytes device-to-device")


def traffic_key(workload, history, layout, schedule, scatter_steps=None, paired=False):
    """Key of a workload's record in profiles/traffic.json."""
    key = f"{workload}:{history}:{schedule}" if layout == "node" else f"{workload}:{history}:{layout}:{schedule}"
    return key + (f":k{scatter_steps}" if scatter_steps and scatter_steps > 1 else "") + (":p2" if paired else "")


def load_traffic(workload, history, layout, schedule, scatter_steps=None, paired=False):
    """HBM bytes per step of the step's kernels, from the rocprofv3 counter runs
    summarised in profiles/traffic.json (None if that workload was not profiled).
    With K steps per scatter pass, a step carries 1/K of the scatter pass and of
    its block's early-edge gathers; with paired row passes, half of a pair's row
    pass and node pass."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    try:
        with open(path) as f:
            rec = json.load(f).get(traffic_key(workload, history, layout, schedule, scatter_steps, paired))
    except (OSError, ValueError):
        return None
    if not rec:
        return None
    byt = lambda v: v["fetch_bytes"] + v["write_bytes"]
    kern = rec["kernels"]
    K = scatter_steps if scatter_steps and scatter_steps > 1 else 1

    def targs(name):
        """Template arguments of a kernel name, e.g. 'lev::scatter_kernel<unsigned short, 1, 4, 2>(...)'."""
        inner = name.split("<", 1)[1].split(">", 1)[0] if "<" in name else ""
        return [t.strip() for t in inner.split(",")]

    total = sum(byt(v) for k, v in kern.items() if "step_kernel" in k or "regroup_kernel" in k)
    # paired row passes: the pair's row pass and node pass, once per two steps
    total += sum(byt(v) for k, v in kern.items() if "pair_kernel" in k or "pair_finish_kernel" in k) / 2
    # the staged schedule's regroup of K steps at once: regroup_multi_kernel<H, K>
    total += sum(byt(v) for k, v in kern.items() if "regroup_multi_kernel" in k and targs(k)[1:2] == [str(K)]) / K
    # scatter_kernel<H, LAYOUT, K, ...>
    scat = [(k, v) for k, v in kern.items() if "scatter_kernel" in k and (K == 1 or targs(k)[2:3] == [str(K)])]
    if scat:
        k, v = scat[0]
        per_block = byt(v)
        early = [e for n, e in kern.items() if "early_kernel" in n]
        if K > 1 and early:
            per_block += byt(early[0]) * early[0]["dispatches"] / v["dispatches"]
        total += per_block / K
    return total or None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps (default 200; the reservoir workload collects 50,000 driven steps)")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="ba10m", choices=["ba10m", "ws100k", "ensemble", "reservoir"])
    ap.add_argument("--n", type=int, default=None, help="override the node count (testing only)")
    ap.add_argument("--layout", default="slot", choices=["node", "slot", "auto"],
                    help="history ring of the f32-tuned and u16 forms (the plain drop-in uses leviathan_init's "
                         "automatic choice)")
    ap.add_argument("--ept", type=int, default=8)
    ap.add_argument("--schedule", default="auto", choices=["auto", "direct", "binned", "staged"])
    ap.add_argument("--bucket-edges", type=int, default=0)
    ap.add_argument("--scatter-steps", default="auto", choices=["auto", "1", "2", "4"],
                    help="steps per scatter pass of the two-phase schedules (default: timed and chosen)")
    ap.add_argument("--row-steps", type=int, default=None, choices=[1, 2],
                    help="steps per row pass of the f32-tuned and u16 forms (default: the engine's, 2 where the "
                         "graph allows paired row passes)")
    ap.add_argument("--static-reservoir", action="store_true", help="reservoir: no plasticity, no controller")
    ap.add_argument("--sweep-alphas", default="", help="reservoir: comma-separated ridge parameters to refit with")
    ap.add_argument("--sweep-washouts", default="", help="reservoir: comma-separated washout lengths to refit with")
    ap.add_argument("--graph-chunk", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-sub", action="store_true",
                    help="skip the sub-records (WS N=10^5 drop-in, 256-network ensemble) of the default line")
    ap.add_argument("--forms", default="f32,f32-tuned,u16",
                    help="f32: plain drop-in (direct/8); f32-tuned, u16: --schedule (default autotuned)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run a plain device copy of BYTES (profiler counter scaling)")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 50_000 if args.workload == "reservoir" else 200
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this build has no multi-GPU "
              "path: the step shards by node ranges, which is described but not run.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        from leviathan_v1_amd import ensemble_bench

        print(json.dumps(ensemble_bench.run(args.steps, args.warmup)))
        return
    if args.workload == "reservoir":
        from leviathan_v1_amd import reservoir_bench

        out = reservoir_bench.run(steps=max(args.steps, 2000), static=args.static_reservoir,
                                  graph_chunk=args.graph_chunk,
                                  sweep_alphas=[float(v) for v in args.sweep_alphas.split(",") if v],
                                  sweep_washouts=[int(v) for v in args.sweep_washouts.split(",") if v])
        print(json.dumps({"metric": "reservoir readout NRMSE (u(t+10)), BF16 MFMA Gram + FP64 refinement",
                          "value": out["nrmse_test"], "unit": "NRMSE", "n_gpus": 1, "steps": max(args.steps, 2000),
                          "higher_is_better": False, "data": "synthetic", **out}))
        return

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    t0 = time.perf_counter()
    wl = make_workload(args.workload, args.n)
    log(f"workload {wl.name}: N={wl.N} E={wl.E} built in {time.perf_counter() - t0:.1f} s")
    forms = {}
    for f in args.forms.split(","):
        if f == "f32":
            # the plain drop-in exactly as leviathan_init configures it
            forms[f] = run_form(wl, "f32", args.steps, args.warmup, "auto", 8, "direct")
        else:
            hist = "f32" if f == "f32-tuned" else f
            forms[f] = run_form(wl, hist, args.steps, args.warmup, args.layout, args.ept, args.schedule,
                                args.bucket_edges, scatter_steps=args.scatter_steps, row_steps=args.row_steps)
        log(f, json.dumps(forms[f]))
    head = forms.get("u16") or forms.get("f32-tuned") or forms["f32"]
    hist = head["history"]
    alg = algorithmic_bytes(wl.N, wl.E, hist)
    # the step's kernels: the fused step kernel (direct) or scatter + step kernels (two-phase schedules;
    # with several steps per scatter pass the scatter time is that of the pass and the early-edge
    # gathers, per step)
    step_ms = head["step_kernel_ms"] + head["scatter_kernel_ms"] + head["regroup_kernel_ms"]
    kernels = {"direct": "step_kernel", "binned": "scatter_kernel+step_kernel",
               "staged": "scatter_kernel+regroup_kernel+step_kernel"}[head["schedule"].split("/")[0]]
    paired = head.get("row_steps") == 2 and (head["schedule"].startswith("direct") or
                                             (head.get("scatter_steps") or 1) % 2 == 0)
    K = head.get("scatter_steps") or 1
    if K > 1:
        # the early-edge gathers: their own kernel, or (staged, paired) the second pair's row pass
        early = "" if paired and head["schedule"].startswith("staged") else "+early_kernel"
        kernels = kernels.replace("scatter_kernel", f"(scatter_kernel{early})/{K}")
        kernels = kernels.replace("regroup_kernel", f"regroup_multi_kernel/{K}")
    if paired:
        kernels = kernels.replace("step_kernel", "(pair_kernel+pair_finish_kernel)/2")
    achieved = alg / (step_ms * 1e-3) / 1e9
    cpu = None if args.no_cpu else cpu_baseline(wl)
    if cpu:
        log("cpu", json.dumps(cpu))
    subs = {}
    if args.workload == "ba10m" and not args.no_sub and args.n is None:
        subs["ws100k_plain_fp32"] = ws100k_record()
        log("ws100k", json.dumps(subs["ws100k_plain_fp32"]))
        from leviathan_v1_amd import ensemble_bench

        ens = ensemble_bench.run(steps=200, warmup=20, seq_steps=5)
        subs["ensemble256"] = {k: ens[k] for k in ("value", "unit", "ms_per_step", "steps", "config",
                                                   "sequential_plain", "speedup_vs_sequential", "step_kernel_ms",
                                                   "r_of_K")}
        log("ensemble", ens["value"], ens["speedup_vs_sequential"])
    out = {
        "metric": METRIC,
        "value": head["edge_steps_per_s"],
        "unit": "edge-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded graph and initial state)",
        "config": {"workload": wl.name, "N": wl.N, "E": wl.E, "max_delay": wl.max_delay,
                   "history": hist, "layout": head["layout"], "schedule": head["schedule"],
                   "scatter_steps": head.get("scatter_steps"), "row_steps": head.get("row_steps")},
        "hbm_gbs": head["hbm_gbs"],
        "plain_fp32": forms.get("f32"),
        "fp32_tuned": forms.get("f32-tuned"),
        "u16": forms.get("u16"),
        "speedup_u16_vs_fp32": (forms["u16"]["edge_steps_per_s"] / forms["f32"]["edge_steps_per_s"]
                                if "u16" in forms and "f32" in forms else None),
        "speedup_u16_vs_fp32_tuned": (forms["u16"]["edge_steps_per_s"] / forms["f32-tuned"]["edge_steps_per_s"]
                                      if "u16" in forms and "f32-tuned" in forms else None),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": load_traffic(args.workload, hist, head["layout"], head["schedule"],
                                             head.get("scatter_steps"), paired),
                     "kernel": kernels, "kernel_ms": step_ms, "algorithmic_bytes_per_launch": alg},
        "cpu_baseline": cpu,
        **subs,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
