"""bench.py's accounting, on the CPU: the algorithmic bytes of a step (the numerator of the
roofline), the keys of the HBM traffic records, the per-step traffic of the multi-step and paired
forms recomputed from the committed counter records, and the one-GPU contract."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

N, E = 10_000_000, 199_999_800  # Barabasi-Albert N = 10^7, m = 10


def test_algorithmic_bytes():
    # per edge: col 4 + delay 1 + weight read and write 8 + the delayed phase (4 | 2);
    # per node: row_ptr 4 + theta, theta_hat, omega_hat read and write 24 + omega 4 + history write (4 | 2)
    assert bench.algorithmic_bytes(N, E, "f32") == 17 * E + 36 * N
    assert bench.algorithmic_bytes(N, E, "u16") == 15 * E + 34 * N
    assert abs(bench.algorithmic_bytes(N, E, "u16") / 1e9 - 3.340) < 1e-3


def test_traffic_keys():
    assert bench.traffic_key("ba10m", "f32", "node", "direct/8") == "ba10m:f32:direct/8"
    assert bench.traffic_key("ba10m", "u16", "slot", "staged/8", 4) == "ba10m:u16:slot:staged/8:k4"
    assert bench.traffic_key("ba10m", "u16", "slot", "staged/8", 4, True) == "ba10m:u16:slot:staged/8:k4:p2"
    assert bench.traffic_key("ba10m", "f32", "node", "direct/8", None, True) == "ba10m:f32:direct/8:p2"
    assert bench.load_traffic("ba10m", "u16", "slot", "no-such-plan") is None


def _records():
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        return json.load(f)


def _bytes(rec, name):
    """Fetch + write bytes per dispatch of the one kernel whose name holds `name`, and its dispatches."""
    hits = [v for k, v in rec["kernels"].items() if name in k]
    assert len(hits) == 1, (name, list(rec["kernels"]))
    return hits[0]["fetch_bytes"] + hits[0]["write_bytes"], hits[0]["dispatches"]


def test_paired_headline_traffic():
    """Four steps per scatter pass with paired row passes: a step carries half of a pair's row
    pass and node pass and a quarter of the scatter and regroup passes (no early-gather kernel)."""
    rec = _records()["ba10m:u16:slot:staged/8:k4:p2"]
    pair, _ = _bytes(rec, "pair_kernel<")
    fin, _ = _bytes(rec, "pair_finish_kernel<")
    sc, _ = _bytes(rec, "scatter_kernel<")
    rg, _ = _bytes(rec, "regroup_multi_kernel<unsigned short, 4")
    want = (pair + fin) / 2 + (sc + rg) / 4
    got = bench.load_traffic("ba10m", "u16", "slot", "staged/8", 4, True)
    assert got == pytest.approx(want, rel=1e-12)
    assert 3.34e9 < got < 5.0e9  # above the algorithmic bytes, below the unpaired form's 4.97 GB


def test_unpaired_multistep_traffic():
    """One row pass per step, four steps per scatter pass: a step carries its row pass, a quarter of
    the scatter and regroup passes, and a quarter of the block's early-gather launches."""
    rec = _records()["ba10m:u16:slot:staged/8:k4"]
    row, _ = _bytes(rec, "step_kernel<")
    sc, n_sc = _bytes(rec, "scatter_kernel<")
    rg, _ = _bytes(rec, "regroup_multi_kernel<unsigned short, 4")
    early, n_early = _bytes(rec, "early_kernel<")
    want = row + rg / 4 + (sc + early * n_early / n_sc) / 4
    assert bench.load_traffic("ba10m", "u16", "slot", "staged/8", 4) == pytest.approx(want, rel=1e-12)


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2 and "one GPU" in p.stderr and p.stdout == ""
