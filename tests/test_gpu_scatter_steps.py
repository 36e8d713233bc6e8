"""Multi-step scatter of the two-phase schedules (leviathan_set_scatter_steps).

One scatter pass gathers the delayed phases of K consecutive steps; step k of the block
first gathers its edges with delay < k, whose phases do not exist at the block's start.
Every value that reaches the row pass is a copy of the same history slot as with K = 1, so
phases, weights, r and the history ring must be bit-identical to K = 1: the bar here is
exact equality, for every two-phase schedule, both ring layouts and both history codings,
on a Barabasi-Albert graph with delays 1..50 (the benchmark's kind), on one whose delays 0..5
make the early edges a large share (delay 0 reads the phases the previous step wrote), and
on one with a source beyond one scatter block.
Step counts that are not multiples of K leave a remainder of single steps; 200 steps cross
captured 100-step graphs.
"""
import dataclasses

import numpy as np
import pytest

from gpu_helpers import say

pytestmark = pytest.mark.gpu

DT = 0.05


@pytest.fixture(scope="module")
def ba50k():
    from leviathan_v1_amd import workloads

    return workloads.barabasi_albert_workload(50_000, 10, 50, seed=8)


@pytest.fixture(scope="module")
def hub6k():
    """WS N=6,000 plus node 0 joined to every other node, delays 0..5: node 0's ~6,000
    out-edges exceed one scatter block (4,096), so the long-source path runs, with early
    edges among them."""
    from leviathan_v1_amd import graph, workloads

    ws = graph.watts_strogatz(6000, 20, 0.1, 12)
    rows = np.repeat(np.arange(ws.num_nodes, dtype=np.int32), ws.degrees())
    up = ws.col_idx > rows
    others = np.arange(1, ws.num_nodes, dtype=np.int32)
    g = graph.from_edges(ws.num_nodes, np.concatenate([rows[up], np.zeros_like(others)]),
                         np.concatenate([ws.col_idx[up], others]))
    assert g.degrees()[0] > 4096
    d, w, th, thh, om = workloads.init_state(g, 5, 13)
    d = np.random.default_rng(14).integers(0, 6, g.num_edges).astype(np.uint8)
    return workloads.Workload("ws6k_hub_d0", g, 5, d, w, th, thh, om)


@pytest.fixture(scope="module")
def ws_d0():
    from leviathan_v1_amd import workloads

    wl = workloads.watts_strogatz_workload(3000, 20, 0.1, 5, seed=31)
    rng = np.random.default_rng(9)
    return dataclasses.replace(wl, delays=rng.integers(0, 6, wl.E).astype(np.uint8))


def make(wl, **kw):
    from leviathan_v1_amd import LeviathanEngine

    return LeviathanEngine(*wl.engine_args(), **kw)


def _same_state(a, b, max_delay):
    assert np.array_equal(a.get_theta(), b.get_theta())
    assert np.array_equal(a.get_weights(), b.get_weights())
    for d in sorted({0, min(1, max_delay), min(2, max_delay), max_delay}):
        assert np.array_equal(a.get_history(d), b.get_history(d)), d


@pytest.mark.parametrize("K", [2, 4])
@pytest.mark.parametrize("schedule", ["binned", "staged"])
@pytest.mark.parametrize("layout", ["slot", "node"])
@pytest.mark.parametrize("history", ["u16", "f32"])
@pytest.mark.parametrize("graph", ["ba50k", "ws_d0", "hub6k"])
def test_multi_step_scatter_bit_identical(request, graph, history, layout, schedule, K):
    wl = request.getfixturevalue(graph)
    say("scatter steps", graph, history, layout, schedule, K)
    a = make(wl, history=history, layout=layout, schedule=schedule)
    b = make(wl, history=history, layout=layout, schedule=schedule)
    b.set_scatter_steps(K)
    assert (a.scatter_steps, b.scatter_steps) == (1, K)
    for n in (37, 200, 1, 6):
        ra = a.step_n(DT, n)
        rb = b.step_n(DT, n)
        assert np.array_equal(ra, rb), n
    rs = [b.step(DT) for _ in range(3)]
    ra = [a.step(DT) for _ in range(3)]
    assert rs == ra
    _same_state(a, b, wl.max_delay)
    a.close()
    b.close()


def test_scatter_steps_autotune_restores_state(ba50k):
    say("scatter steps autotune")
    a = make(ba50k, history="u16", schedule="binned")
    ref = make(ba50k, history="u16", schedule="binned")
    a.step_n(DT, 7)
    ref.step_n(DT, 7)
    times = a.autotune_scatter_steps(DT, steps=8)
    say("ms/step by scatter steps", times)
    assert set(times) == {1, 2, 4}
    assert a.scatter_steps == min(times, key=times.get)
    assert a.step_count == 7
    _same_state(a, ref, ba50k.max_delay)
    # profile_steps runs whole blocks too and leaves the same state
    prof = a.profile_steps(DT, 12)
    ref.step_n(DT, 12)
    assert prof["scatter_kernel_ms"] > 0.0 and prof["regroup_kernel_ms"] == 0.0
    _same_state(a, ref, ba50k.max_delay)
    np.testing.assert_array_equal(a.step_n(DT, 20), ref.step_n(DT, 20))
    # the direct schedule has no scatter: nothing is timed or changed
    d = make(ba50k)
    assert d.autotune_scatter_steps(DT, steps=4) == {}
    with pytest.raises(RuntimeError, match="two-phase"):
        d.set_scatter_steps(2)
    with pytest.raises(RuntimeError, match="1, 2 or 4"):
        a.set_scatter_steps(3)
    for x in (a, ref, d):
        x.close()


@pytest.mark.parametrize("schedule", ["binned", "staged"])
def test_autotune_scatter_steps_after_captured_graphs(ba50k, schedule):
    """step_n captures 100-step graphs that hold the two-phase value arrays' addresses.  The
    K autotune grows those arrays (room for 4 steps) whatever K it keeps, so every captured
    graph must be dropped with the old arrays: the next step_n replays graphs of the new
    arrays and continues exactly like an engine that never autotuned."""
    say("scatter steps autotune after captured graphs", schedule)
    a = make(ba50k, history="u16", schedule=schedule)
    ref = make(ba50k, history="u16", schedule=schedule)
    np.testing.assert_array_equal(a.step_n(DT, 200), ref.step_n(DT, 200))  # two captured graphs each
    a.autotune_scatter_steps(DT, steps=8)
    ref.set_scatter_steps(a.scatter_steps)
    say("kept K =", a.scatter_steps)
    for n in (200, 100, 3):
        np.testing.assert_array_equal(a.step_n(DT, n), ref.step_n(DT, n))
    _same_state(a, ref, ba50k.max_delay)
    a.close()
    ref.close()


def test_checkpoint_across_scatter_steps(ws_d0, tmp_path):
    """A checkpoint taken mid-run with K = 4 restores into an engine with K = 2 and continues
    bit for bit like the original."""
    from leviathan_v1_amd import checkpoint

    say("scatter steps checkpoint")
    a = make(ws_d0, history="u16", schedule="staged")
    a.set_scatter_steps(4)
    a.step_n(DT, 21)
    path = str(tmp_path / "ck.npz")
    checkpoint.save(path, a.checkpoint())
    b = make(ws_d0, history="u16", schedule="staged")
    b.set_scatter_steps(2)
    b.restore(checkpoint.load(path))
    np.testing.assert_array_equal(a.step_n(DT, 30), b.step_n(DT, 30))
    _same_state(a, b, ws_d0.max_delay)
    a.close()
    b.close()


@pytest.mark.parametrize("max_delay", [0, 1, 2])
@pytest.mark.parametrize("schedule", ["binned", "staged"])
def test_short_rings(max_delay, schedule):
    """Rings of S = max_delay + 2 slots take at most S steps per pass; with delays 0..max_delay
    every edge of every later step of a block is early."""
    from leviathan_v1_amd import workloads

    say("scatter steps short ring", max_delay, schedule)
    wl = workloads.watts_strogatz_workload(2000, 20, 0.1, max(max_delay, 1), seed=40 + max_delay)
    wl = dataclasses.replace(wl, max_delay=max_delay,
                             delays=np.random.default_rng(max_delay).integers(0, max_delay + 1, wl.E)
                             .astype(np.uint8))
    for K in (2, 4):
        b = make(wl, history="u16", schedule=schedule)
        if K > max_delay + 2:
            with pytest.raises(RuntimeError, match="exceed the ring"):
                b.set_scatter_steps(K)
            assert set(b.autotune_scatter_steps(DT, steps=4)) == {k for k in (1, 2, 4) if k <= max_delay + 2}
            b.close()
            continue
        b.set_scatter_steps(K)
        ref = make(wl, history="u16", schedule=schedule)
        np.testing.assert_array_equal(ref.step_n(DT, 45), b.step_n(DT, 45))
        _same_state(ref, b, max_delay)
        ref.close()
        b.close()


@pytest.mark.parametrize("schedule", ["binned", "staged"])
def test_ensemble_and_driven_feature_capture(schedule):
    """An ensemble of 12 networks (the finalize grid) with the reservoir drive and feature capture
    on: several steps per pass give the same r of every network, the same phases and the same
    captured feature rows as one step per pass."""
    from leviathan_v1_amd import LeviathanEngine, graph, workloads

    say("scatter steps ensemble + features", schedule)
    wls = [workloads.watts_strogatz_workload(400 + 30 * k, 20, 0.1, 6, seed=70 + k) for k in range(12)]
    g, offs = graph.disjoint_union([w.graph for w in wls])
    cat = lambda f: np.concatenate([getattr(w, f) for w in wls])
    delays = np.random.default_rng(3).integers(0, 7, g.num_edges).astype(np.uint8)
    args = (g.num_nodes, 6, g.row_ptr, g.col_idx, delays, cat("weights"), cat("theta"), cat("theta_hat"),
            cat("omega"))
    kw = dict(n_nets=12, net_offsets=offs, g0=list(np.linspace(0, 4, 12)), controller=[k % 2 == 0 for k in range(12)],
              history="u16", schedule=schedule)
    rng = np.random.default_rng(4)
    drive_b = rng.normal(0, 0.3, g.num_nodes).astype(np.float32)
    u = rng.normal(0, 1, 230).astype(np.float32)
    nodes = np.sort(rng.choice(g.num_nodes, 64, replace=False)).astype(np.int32)
    engines = []
    for K in (1, 4):
        e = LeviathanEngine(*args, **kw)
        if K > 1:
            e.set_scatter_steps(K)
        e.set_drive(drive_b)
        e.features_setup(nodes, 300)
        engines.append((e, e.step_n_features(DT, u=u)))
    (a, ra), (b, rb) = engines
    np.testing.assert_array_equal(ra, rb)
    np.testing.assert_array_equal(a.features_get(0, 230), b.features_get(0, 230))
    assert np.array_equal(a.get_theta(), b.get_theta())
    a.close()
    b.close()


def test_autotune_all_restores_state(ba50k):
    """Plans, steps per pass and plans again, each from the same state: the state is restored,
    and the chosen configuration steps like a direct engine (every plan and every K gives the
    same phases and weights)."""
    say("autotune_all")
    a = make(ba50k, history="u16", schedule="auto")
    ref = make(ba50k, history="u16")
    a.step_n(DT, 6)
    ref.step_n(DT, 6)
    res = a.autotune_all(DT, steps=4)
    say("autotune_all", res, a.schedule_name, a.scatter_steps)
    assert len(res["plans"]) == 9 and set(res["scatter_steps"]) == {1, 2, 4}
    assert a.step_count == 6
    _same_state(a, ref, ba50k.max_delay)
    a.step_n(DT, 24)
    ref.step_n(DT, 24)
    assert np.array_equal(a.get_weights(), ref.get_weights())
    assert np.array_equal(a.get_theta(), ref.get_theta())
    a.close()
    ref.close()
