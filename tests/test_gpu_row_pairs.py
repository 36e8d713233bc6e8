"""Paired row passes (leviathan_set_row_steps).

Without delay-0 edges, step t+1 reads no phase that step t writes.  So one row pass runs step t
in full and step t+1's prediction, Hebbian update and coupling sum (each weight read and written
once for the two steps), and a node pass integrates step t+1 once step t's finalize has set the
new gain.  Every value is computed from the same operands in the same order as two single steps,
so the bar here is exact equality of r, phases, weights, predictor, history and gains against
row_steps = 1, for every schedule, both ring layouts and both history codings (the staged
schedule with uint16 history regroups the values of each pair of steps side by side, one 4-byte
pair per slot, where single steps get one array per step), on graphs with
  * rows beyond one block (long-row mode: two sweeps over the row, each lane reading back only
    the weights it stored),
  * many short delays (delays 1..5: many early edges, which the second pair of a staged block
    gathers itself and the binned schedule gathers before the pair, the second step's a step
    ahead),
  * a hub whose ~6,000 out-edges exceed one scatter block.
Odd step counts leave single steps; 200 steps cross captured 100-step graphs.
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
def ws_d1():
    from leviathan_v1_amd import workloads

    wl = workloads.watts_strogatz_workload(3000, 20, 0.1, 5, seed=31)
    rng = np.random.default_rng(19)
    return dataclasses.replace(wl, delays=rng.integers(1, 6, wl.E).astype(np.uint8))


@pytest.fixture(scope="module")
def hub6k_d1():
    """WS N=6,000 plus node 0 joined to every other node (in both directions), delays 1..5:
    node 0's ~6,000 incoming edges make a long row, its out-edges a long scatter source."""
    from leviathan_v1_amd import graph, workloads

    ws = graph.watts_strogatz(6000, 20, 0.1, 12)
    rows = np.repeat(np.arange(ws.num_nodes, dtype=np.int32), ws.degrees())
    up = ws.col_idx > rows
    others = np.arange(1, ws.num_nodes, dtype=np.int32)
    g = graph.from_edges(ws.num_nodes, np.concatenate([rows[up], np.zeros_like(others)]),
                         np.concatenate([ws.col_idx[up], others]))
    assert g.degrees()[0] > 4096
    _, w, th, thh, om = workloads.init_state(g, 5, 13)
    d = np.random.default_rng(14).integers(1, 6, g.num_edges).astype(np.uint8)
    return workloads.Workload("ws6k_hub_d1", g, 5, d, w, th, thh, om)


def make(wl, **kw):
    from leviathan_v1_amd import LeviathanEngine

    return LeviathanEngine(*wl.engine_args(), **kw)


def _same_state(a, b, max_delay, n_nets=1):
    assert np.array_equal(a.get_theta(), b.get_theta())
    assert np.array_equal(a.get_weights(), b.get_weights())
    pa, pb = a.get_predictor(), b.get_predictor()
    assert np.array_equal(pa[0], pb[0]) and np.array_equal(pa[1], pb[1])
    assert np.array_equal(a.get_gains(), b.get_gains())
    for d in sorted({1, min(2, max_delay), max_delay}):
        assert np.array_equal(a.get_history(d), b.get_history(d)), d


PLANS = [("direct", 1), ("binned", 2), ("binned", 4), ("staged", 2), ("staged", 4)]


@pytest.mark.parametrize("schedule,K", PLANS)
@pytest.mark.parametrize("layout", ["slot", "node"])
@pytest.mark.parametrize("history", ["u16", "f32"])
@pytest.mark.parametrize("graph", ["ba50k", "ws_d1", "hub6k_d1"])
def test_paired_row_pass_bit_identical(request, graph, history, layout, schedule, K):
    wl = request.getfixturevalue(graph)
    say("row pairs", graph, history, layout, schedule, K)
    # four edges per lane: blocks of 1,024 edges, so the BA graph's hubs are long rows too
    kw = dict(history=history, layout=layout, schedule=schedule, edges_per_thread=4)
    a = make(wl, **kw)
    b = make(wl, **kw)
    assert (a.row_steps, b.row_steps) == (2, 2)  # the default for graphs without delay-0 edges
    a.set_row_steps(1)
    if K > 1:
        a.set_scatter_steps(K)
        b.set_scatter_steps(K)
    for n in (37, 200, 1, 6):
        ra = a.step_n(DT, n)
        rb = b.step_n(DT, n)
        assert np.array_equal(ra, rb), n
    assert [a.step(DT) for _ in range(3)] == [b.step(DT) for _ in range(3)]
    _same_state(a, b, wl.max_delay)
    a.close()
    b.close()


@pytest.mark.parametrize("schedule,K", [("direct", 1), ("staged", 4)])
def test_paired_ensemble_driven_feature_capture(schedule, K):
    """12 networks (the finalize grid, per-network gains, half with the controller on), the
    reservoir drive (the node pass reads the drive of its own step) and feature capture (rows
    written after each step's finalize): paired passes give the same r of every network, the
    same gains, phases and captured feature rows as single steps."""
    from leviathan_v1_amd import LeviathanEngine, graph, workloads

    say("row pairs ensemble + features", schedule, K)
    wls = [workloads.watts_strogatz_workload(400 + 30 * k, 20, 0.1, 6, seed=70 + k) for k in range(12)]
    g, offs = graph.disjoint_union([w.graph for w in wls])
    cat = lambda f: np.concatenate([getattr(w, f) for w in wls])
    delays = np.random.default_rng(3).integers(1, 7, g.num_edges).astype(np.uint8)
    args = (g.num_nodes, 6, g.row_ptr, g.col_idx, delays, cat("weights"), cat("theta"), cat("theta_hat"),
            cat("omega"))
    kw = dict(n_nets=12, net_offsets=offs, g0=list(np.linspace(0, 4, 12)), controller=[k % 2 == 0 for k in range(12)],
              history="u16", schedule=schedule)
    rng = np.random.default_rng(4)
    drive_b = rng.normal(0, 0.3, g.num_nodes).astype(np.float32)
    u = rng.normal(0, 1, 230).astype(np.float32)
    nodes = np.sort(rng.choice(g.num_nodes, 64, replace=False)).astype(np.int32)
    engines = []
    for rows in (1, 2):
        e = LeviathanEngine(*args, **kw)
        e.set_row_steps(rows)
        if K > 1:
            e.set_scatter_steps(K)
        e.set_drive(drive_b)
        e.features_setup(nodes, 300)
        engines.append((e, e.step_n_features(DT, u=u)))
    (a, ra), (b, rb) = engines
    np.testing.assert_array_equal(ra, rb)
    np.testing.assert_array_equal(a.features_get(0, 230), b.features_get(0, 230))
    _same_state(a, b, 6)
    a.close()
    b.close()


def test_row_steps_setting(ba50k):
    """A graph with delay-0 edges refuses paired passes; profile_steps (pass and node pass timed
    as the two steps of a pair) and the autotune leave the state a plain step_n leaves; switching
    the setting between runs keeps the results exact."""
    from leviathan_v1_amd import workloads

    say("row steps setting")
    wl0 = workloads.watts_strogatz_workload(1000, 10, 0.1, 4, seed=3)
    wl0 = dataclasses.replace(wl0, delays=np.random.default_rng(2).integers(0, 5, wl0.E).astype(np.uint8))
    z = make(wl0)
    assert z.row_steps == 1
    with pytest.raises(RuntimeError, match="delay-0"):
        z.set_row_steps(2)
    with pytest.raises(RuntimeError, match="1 or 2"):
        z.set_row_steps(3)
    z.close()

    a = make(ba50k, history="u16", schedule="staged")
    ref = make(ba50k, history="u16", schedule="staged")
    ref.set_row_steps(1)
    for e in (a, ref):
        e.set_scatter_steps(4)
    prof = a.profile_steps(DT, 12)
    say("profile", prof)
    assert prof["step_kernel_ms"] > 0 and prof["scatter_kernel_ms"] > 0 and prof["regroup_kernel_ms"] > 0
    ref.step_n(DT, 12)
    _same_state(a, ref, ba50k.max_delay)
    a.autotune_all(DT, steps=8)
    say("autotuned", a.schedule_name, a.scatter_steps, a.row_steps)
    assert a.step_count == ref.step_count == 12
    _same_state(a, ref, ba50k.max_delay)
    a.set_schedule("staged", 8)
    a.set_scatter_steps(4)
    for rows, n in ((1, 50), (2, 101), (1, 3), (2, 100)):
        a.set_row_steps(rows)
        np.testing.assert_array_equal(a.step_n(DT, n), ref.step_n(DT, n))
    _same_state(a, ref, ba50k.max_delay)
    a.close()
    ref.close()


def test_automatic_layout(ba50k, ws_d1):
    """The default ring layout: node-major only for the direct schedule on a graph without
    delay-0 edges and with little index locality (Barabasi-Albert); slot-major for a ring lattice
    (most sources within 256 nodes), for delay-0 edges, and for two-phase schedules.  An explicit
    layout is kept, and results do not depend on the layout."""
    say("automatic layout")
    ba0 = dataclasses.replace(ba50k, delays=np.where(np.arange(ba50k.E) % 97 == 0, 0, ba50k.delays)
                              .astype(np.uint8))
    cases = [(ba50k, {}, "node"), (ws_d1, {}, "slot"), (ba0, {}, "slot"),
             (ba50k, dict(schedule="staged"), "slot"), (ba50k, dict(schedule="auto"), "slot"),
             (ba50k, dict(layout="slot"), "slot"), (ws_d1, dict(layout="node"), "node")]
    for wl, kw, want in cases:
        e = make(wl, **kw)
        assert e.layout == want, (wl.name, kw)
        e.close()
    a, b = make(ba50k), make(ba50k, layout="slot")
    np.testing.assert_array_equal(a.step_n(DT, 30), b.step_n(DT, 30))
    assert np.array_equal(a.get_theta(), b.get_theta()) and np.array_equal(a.get_weights(), b.get_weights())
    a.close()
    b.close()
