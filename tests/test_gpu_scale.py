"""The full-size step: Barabasi-Albert N = 10^7, m = 10 (2.0e8 incoming edges, in-degrees
up to ~10^4), delays uniform in [1, 50], FP32 and uint16 history.

  * per-step phases and weights of each form within 1e-5 relative of one FP64 oracle step
    taken from the engine's exported state (the bar of the other GPU tests), and the
    step's r within 1e-6 of the FP64 order parameter of the phases it wrote;
  * the benchmark's form (uint16 history, staged schedule with 8 edges per lane, two steps
    per scatter pass, a paired row pass): each step of a two-step block against an FP64
    oracle step from the engine's state, and the block bit-identical to the same two steps
    taken one scatter pass and one row pass each;
  * 2,000 steps of the uint16 form and of the plain FP32 drop-in against the FP64 oracle's
    2,000 steps of the same inputs (the committed fixture tests/golden/ba10m_fp64.npz,
    written by tests/golden/make_golden.py with the OpenMP FP64 oracle): windowed
    (200-step) mean r and r variance within 5e-3.
"""
import hashlib
import os

import numpy as np
import pytest

from gpu_helpers import oracle_from_engine, say, theta_rel_err, w_rel_err, windowed

pytestmark = pytest.mark.gpu

DT = 0.05
TOL = 1e-5
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ba10m_fp64.npz")


@pytest.fixture(scope="module")
def ba10m():
    from leviathan_v1_amd import workloads

    return workloads.barabasi_albert_workload(10_000_000, 10, 50, seed=0)


@pytest.fixture(scope="module")
def golden10m(ba10m):
    assert os.path.exists(GOLDEN), "tests/golden/ba10m_fp64.npz is missing (tests/golden/make_golden.py ba10m)"
    z = np.load(GOLDEN)
    h = hashlib.sha256()
    wl = ba10m
    for a in (wl.graph.row_ptr, wl.graph.col_idx, wl.delays, wl.weights, wl.theta, wl.theta_hat, wl.omega):
        h.update(np.ascontiguousarray(a).tobytes())
    assert str(z["digest"]) == h.hexdigest(), "the fixture was written from other inputs"
    return z


def make(wl, **kw):
    from leviathan_v1_amd import LeviathanEngine

    return LeviathanEngine(*wl.engine_args(), **kw)


@pytest.mark.parametrize("history,schedule", [("f32", "direct"), ("u16", "binned"), ("u16", "staged")])
def test_ba10m_one_step_vs_fp64_oracle(ba10m, history, schedule):
    say("ba10m one step", history, schedule)
    eng = make(ba10m, history=history, schedule=schedule)
    eng.step_n(DT, 20)
    o = oracle_from_engine(eng, ba10m, parallel=True)
    o.step(DT)
    r = eng.step(DT)
    th = eng.get_theta()
    et, ew = theta_rel_err(th, o.theta), w_rel_err(eng.get_weights(), o.w)
    # r of the step (10^5 row-block partials, summed in chunks, then the chunk sums) against
    # the FP64 order parameter of the phases the step wrote
    r_np = float(np.abs(np.exp(1j * th.astype(np.float64)).mean()))
    say("theta rel err", et, "w rel err", ew, "r", r, "r of phases", r_np, "blocks", eng.num_blocks)
    eng.close()
    assert et < TOL and ew < TOL
    assert abs(r - r_np) < 1e-6


def test_ba10m_two_step_block_vs_fp64_oracle(ba10m):
    """The benchmark's form takes its steps in blocks: one scatter pass and one regroup pass
    gather the delayed phases of two steps, then one paired row pass runs the first step and
    the second step's edge work, and a node pass the second step's integration.  From the state at a block
    boundary, each step of the block is compared with an FP64 oracle step taken from the
    engine's own state before it (1e-5 relative): the first from the block engine's exported
    state, the second from the state of an engine that took the first step alone, which is
    bit-identical to the block's (the block's end state is checked bit for bit against it).
    Two free-running FP64 steps are not the bar: the heaviest rows (~15,000 incoming edges,
    coupling derivative ~350) multiply any difference after the first step by ~27 in the
    second, so that comparison measures the dynamics' sensitivity, not the kernel.  It is
    printed for reference."""
    say("ba10m two-step block, u16 staged/8, 2 steps per pass")
    a = make(ba10m, history="u16", schedule="staged", edges_per_thread=8)
    b = make(ba10m, history="u16", schedule="staged", edges_per_thread=8)
    a.set_scatter_steps(2)
    assert (a.scatter_steps, b.scatter_steps) == (2, 1) and a.row_steps == 2
    np.testing.assert_array_equal(a.step_n(DT, 20), b.step_n(DT, 20))
    o1 = oracle_from_engine(a, ba10m, parallel=True)
    o1.step(DT)
    b.step(DT)  # the block's first step, alone
    e1t, e1w = theta_rel_err(b.get_theta(), o1.theta), w_rel_err(b.get_weights(), o1.w)
    o2 = oracle_from_engine(b, ba10m, parallel=True)
    o2.step(DT)
    o1.step(DT)  # free-running second FP64 step, for reference
    b.step(DT)
    ra = a.step_n(DT, 2)  # one block: scatter + regroup of two steps, a paired row pass, a node pass
    th = a.get_theta()
    e2t, e2w = theta_rel_err(th, o2.theta), w_rel_err(a.get_weights(), o2.w)
    say("step 1: theta", e1t, "w", e1w, "| step 2 from the engine's state: theta", e2t, "w", e2w,
        "| free-running FP64 two steps: theta", theta_rel_err(th, o1.theta), "| r", float(ra[-1]))
    assert e1t < TOL and e1w < TOL
    assert e2t < TOL and e2w < TOL
    assert np.array_equal(th, b.get_theta())
    assert np.array_equal(a.get_weights(), b.get_weights())
    for d in (0, 1, 2, 50):
        assert np.array_equal(a.get_history(d), b.get_history(d)), d
    a.close()
    b.close()


@pytest.mark.parametrize("form", ["u16_staged8_k2", "u16_staged8_k4", "f32_plain"])
def test_ba10m_windowed_r_vs_fp64_fixture(ba10m, golden10m, form):
    """2,000 steps against the FP64 oracle's 2,000 steps of the same inputs.  The dynamics are
    chaotic, so trajectories separate; the bar is on windowed (200-step) mean r and r
    variance, within 5e-3 in every window."""
    say("ba10m 2000 steps vs FP64 fixture", form)
    if form.startswith("u16"):
        eng = make(ba10m, history="u16", schedule="staged", edges_per_thread=8)
        eng.set_scatter_steps(int(form[-1]))
    else:
        eng = make(ba10m)  # the plain FP32 drop-in, as leviathan_init configures it
    r = eng.step_n(DT, 2000)
    g = eng.get_gain()
    eng.close()
    ref = golden10m["r"]
    assert np.isfinite(r).all() and 0 <= r.min() and r.max() <= 1
    m, v = windowed(r, 0, 200)
    mr, vr = windowed(ref, 0, 200)
    say(form, "max |mean r diff|", float(np.abs(m - mr).max()), "max |var diff|", float(np.abs(v - vr).max()),
        "g", g, "g fp64", float(golden10m["g_final"]))
    assert np.abs(m - mr).max() < 5e-3
    assert np.abs(v - vr).max() < 5e-3
