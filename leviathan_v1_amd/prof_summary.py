"""Summaries of rocprofv3 output for the step's kernels.

    python -m leviathan_v1_amd.prof_summary stats KERNEL_STATS_CSV --cmd "..." [--top 8]
    python -m leviathan_v1_amd.prof_summary traffic FETCH_CSV WRITE_CSV --copy-bytes B --key KEY \
        --cmd-fetch "..." --cmd-write "..." [--json profiles/traffic.json]

``traffic`` scales each counter by what it reads for the calibration copy of
B bytes (the device-to-device copies bench.py --copy-calibration makes): a
copy of B bytes fetches B and writes B, so scale = B / counter(copy).  It
prints the per-kernel HBM bytes per dispatch and records them in the JSON file
that bench.py reads for ``roofline.traffic``.
"""
from __future__ import annotations

import argparse
import csv
import json
import os
import sys
from collections import defaultdict


def _short(name: str) -> str:
    return (name.replace("(lev::StepArgs)", "").replace("(lev::ScatterArgs)", "").replace("(lev::EarlyArgs)", "")
            .replace("(lev::RegroupArgs)", "").replace("void ", "").strip())


def stats(path, cmd, top=8):
    rows = list(csv.DictReader(open(path)))
    total = sum(float(r["TotalDurationNs"]) for r in rows) or 1.0
    out = [f"command: `{cmd}`", "", "| kernel | calls | avg ms | min ms | max ms | % time |", "|---|---|---|---|---|---|"]
    for r in rows[:top]:
        out.append(f"| {_short(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e6:.4f} | "
                   f"{float(r['MinNs']) / 1e6:.4f} | {float(r['MaxNs']) / 1e6:.4f} | "
                   f"{100 * float(r['TotalDurationNs']) / total:.1f} |")
    return "\n".join(out)


def _per_kernel(path):
    agg = defaultdict(list)
    copies = []
    for r in csv.DictReader(open(path)):
        v = float(r["Counter_Value"])
        if "copyBuffer" in r["Kernel_Name"]:
            copies.append(v)
        agg[_short(r["Kernel_Name"])].append(v)
    return agg, copies


def traffic(fetch_csv, write_csv, copy_bytes, key, cmd_fetch, cmd_write, json_path):
    fa, fc = _per_kernel(fetch_csv)
    wa, wc = _per_kernel(write_csv)
    big = lambda c: [x for x in c if x * 1024 > copy_bytes / 8]  # the calibration copies, not small blits
    fcal, wcal = big(fc), big(wc)
    if not fcal or not wcal:
        raise SystemExit("no calibration copy found: run with --copy-calibration")
    f_scale = copy_bytes / (sum(fcal) / len(fcal) * 1024)
    w_scale = copy_bytes / (sum(wcal) / len(wcal) * 1024)
    lines = [f"fetch command: `{cmd_fetch}`", f"write command: `{cmd_write}`", "",
             f"calibration: {copy_bytes} B device copy reads FETCH_SIZE {sum(fcal) / len(fcal):.0f} KiB, "
             f"WRITE_SIZE {sum(wcal) / len(wcal):.0f} KiB -> scale fetch x{f_scale:.3f}, write x{w_scale:.3f}", "",
             "| kernel | dispatches | HBM fetch GB/dispatch | HBM write GB/dispatch |", "|---|---|---|---|"]
    rec = {}
    for k in fa:
        if not k.startswith("lev::"):
            continue
        fv = sum(fa[k]) / len(fa[k]) * 1024 * f_scale
        wv = sum(wa.get(k, [0.0])) / max(len(wa.get(k, [1])), 1) * 1024 * w_scale
        lines.append(f"| {k} | {len(fa[k])} | {fv / 1e9:.3f} | {wv / 1e9:.3f} |")
        rec[k] = {"fetch_bytes": fv, "write_bytes": wv, "dispatches": len(fa[k])}
    db = {}
    if os.path.exists(json_path):
        with open(json_path) as f:
            db = json.load(f)
    db[key] = {"kernels": rec, "fetch_scale": f_scale, "write_scale": w_scale,
               "commands": [cmd_fetch, cmd_write]}
    with open(json_path, "w") as f:
        json.dump(db, f, indent=1, sort_keys=True)
    return "\n".join(lines)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--cmd", required=True)
    s.add_argument("--top", type=int, default=8)
    t = sub.add_parser("traffic")
    t.add_argument("fetch_csv")
    t.add_argument("write_csv")
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--key", required=True)
    t.add_argument("--cmd-fetch", required=True)
    t.add_argument("--cmd-write", required=True)
    t.add_argument("--json", default="profiles/traffic.json")
    a = ap.parse_args(argv)
    if a.what == "stats":
        print(stats(a.csv, a.cmd, a.top))
    else:
        print(traffic(a.fetch_csv, a.write_csv, a.copy_bytes, a.key, a.cmd_fetch, a.cmd_write, a.json))


if __name__ == "__main__":
    sys.exit(main())
