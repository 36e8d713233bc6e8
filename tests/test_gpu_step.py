"""GPU parity of the fused gfx950 step against the FP64 oracle.

Tolerances (stated once, used below):
  * per-step phases and weights within 1e-5 relative: phase error on the circle
    / 2*pi, and |w - w_ref| / max(|w_ref|, 1e-3);
  * trajectories from the same inputs: the 1e-5 bar holds for 10 steps (the
    dynamics are chaotic; FP32 rounding differences grow, see test_oracle);
  * one-step consistency: from the GPU's exported state at any time, one
    oracle step and one GPU step agree within 1e-5;
  * long runs: windowed mean r and r variance within 5e-3 of the oracle.
"""
import os

import numpy as np
import pytest

from gpu_helpers import oracle_from_engine, say, theta_rel_err, w_rel_err, windowed

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DT = 0.05
TOL = 1e-5


@pytest.fixture(scope="module")
def ws1k():
    from leviathan_v1_amd import workloads

    return workloads.small_plain(0)


@pytest.fixture(scope="module")
def golden1k():
    return np.load(os.path.join(HERE, "golden", "ws1k_fp64.npz"))


def make(wl, **kw):
    from leviathan_v1_amd import LeviathanEngine

    return LeviathanEngine(*wl.engine_args(), **kw)


CONFIGS = [
    dict(),
    dict(layout="slot"),
    dict(edges_per_thread=4),
    dict(edges_per_thread=16),
    dict(graph_chunk=0),
]


@pytest.mark.parametrize("kw", CONFIGS, ids=lambda k: ",".join(f"{a}={b}" for a, b in k.items()) or "default")
def test_trajectory_vs_fp64_golden(ws1k, golden1k, kw):
    """WS N=1,000: theta and W at steps 1 and 10 against the FP64 golden vectors."""
    say("ws1k trajectory", kw)
    from tests.golden.make_golden import digest

    assert str(golden1k["digest"]) == digest(ws1k), "workload generator changed"
    eng = make(ws1k, **kw)
    r = [eng.step(DT)]
    assert theta_rel_err(eng.get_theta(), golden1k["theta_1"]) < TOL
    assert w_rel_err(eng.get_weights(), golden1k["w_1"]) < TOL
    r += list(eng.step_n(DT, 9))
    assert theta_rel_err(eng.get_theta(), golden1k["theta_10"]) < TOL
    assert w_rel_err(eng.get_weights(), golden1k["w_10"]) < TOL
    assert np.abs(np.array(r) - golden1k["r"][:10]).max() < TOL
    assert abs(eng.get_gain() - float(golden1k["g_10"])) < TOL


def test_long_run_r_statistics_ws1k(ws1k, golden1k):
    """500 steps (the whole small workload): windowed r mean/variance within 5e-3."""
    say("ws1k 500 steps r statistics")
    eng = make(ws1k)
    r = eng.step_n(DT, 500)
    m, v = windowed(r, 100, 100)
    mr, vr = windowed(golden1k["r"], 100, 100)
    assert np.abs(m - mr).max() < 5e-3
    assert np.abs(v - vr).max() < 5e-3
    assert 0.0 <= r.min() and r.max() <= 1.0
    th = eng.get_theta()
    assert np.isfinite(th).all() and th.min() >= 0 and th.max() <= 2 * np.pi + 1e-5


@pytest.mark.parametrize("history", ["f32", "u16"])
@pytest.mark.parametrize("layout", ["node", "slot"])
def test_one_step_consistency(ws1k, history, layout):
    """At steps 0, 57 and 300, one GPU step equals one FP64 oracle step from the GPU's own state."""
    say("one-step consistency", history, layout)
    eng = make(ws1k, history=history, layout=layout)
    done = 0
    for t in (0, 57, 300):
        if t > done:
            eng.step_n(DT, t - done)
            done = t
        o = oracle_from_engine(eng, ws1k)
        r_ref = o.step(DT)
        r = eng.step(DT)
        done += 1
        assert theta_rel_err(eng.get_theta(), o.theta) < TOL, t
        assert w_rel_err(eng.get_weights(), o.w) < TOL, t
        th, om = eng.get_predictor()
        assert np.abs(th - o.theta_hat).max() < 1e-4 and np.abs(om - o.omega_hat).max() < 1e-5
        assert abs(r - r_ref) < TOL
        assert abs(eng.get_gain() - o.g) < TOL


def test_u16_history_error_vs_fp64(ws1k, golden1k):
    """uint16 exact-wrap history: stored phases are within pi/65536 of the FP32 phase,
    the 10-step trajectory stays within 1e-3 rad of FP64, and long-run r statistics
    within 5e-3."""
    say("u16 error bounds")
    eng = make(ws1k, history="u16")
    eng.step_n(DT, 3)
    th = eng.get_theta()
    q = eng.get_history(0)
    err = np.abs((q.astype(np.float64) - th + np.pi) % (2 * np.pi) - np.pi)
    assert err.max() <= np.pi / 65536 * 1.01
    eng2 = make(ws1k, history="u16")
    r1 = eng2.step_n(DT, 10)
    e10 = float(np.abs((eng2.get_theta() - golden1k["theta_10"] + np.pi) % (2 * np.pi) - np.pi).max())
    say("u16 10-step max phase error vs FP64 (rad):", e10)
    assert e10 < 1e-3
    r = np.concatenate([r1, eng2.step_n(DT, 490)])
    m, v = windowed(r, 100, 100)
    mr, vr = windowed(golden1k["r"], 100, 100)
    assert np.abs(m - mr).max() < 5e-3 and np.abs(v - vr).max() < 5e-3


def test_step_n_graphs_equal_single_steps(ws1k):
    """250 steps as step() calls and as step_n (two captured 100-step graphs + 50 eager) agree bitwise."""
    say("step_n vs step")
    a = make(ws1k)
    b = make(ws1k)
    ra = np.array([a.step(DT) for _ in range(250)], np.float32)
    rb = b.step_n(DT, 250)
    assert np.array_equal(ra, rb)
    assert np.array_equal(a.get_theta(), b.get_theta())
    assert np.array_equal(a.get_weights(), b.get_weights())
    assert a.get_gain() == b.get_gain()
    assert a.step_count == b.step_count == 250


def test_graph_replays_stay_correct(ws1k):
    """Replaying the same captured graph many times, with eager work between replays."""
    say("graph replays")
    a = make(ws1k, graph_chunk=10)
    b = make(ws1k, graph_chunk=0)
    for k in range(6):
        ra = a.step_n(DT, 30)
        rb = b.step_n(DT, 30)
        assert np.array_equal(ra, rb), k
        a.step(DT)
        b.step(DT)
        assert np.array_equal(a.get_theta(), b.get_theta()), k


@pytest.mark.parametrize("d", [1, 7, 12])
def test_delay_semantics_two_nodes(d):
    """Edge 0->1 with delay d (max_delay 12).  Node 0 rotates freely; history(k) is
    theta_0(t-k) (theta_0(0) before time 0), and row 1 moves by
    dt*g*w*sin(theta_0(t-d) - theta_1(t)) -- d = 12 reads the oldest slot."""
    from leviathan_v1_amd import LeviathanEngine

    say("delay semantics", d)
    D = 12
    th0 = np.array([0.4, 2.0], np.float32)
    om = np.array([1.1, 0.0], np.float32)
    w, g = 0.3, 1.0
    eng = LeviathanEngine(2, D, np.array([0, 0, 1], np.int32), np.array([0], np.int32), np.array([d], np.uint8),
                          np.array([w], np.float32), th0, th0.copy(), om,
                          params=dict(g0=g, gain_controller=False, alpha_eps=0.0, eta=0.0, lam=0.0))
    circ = lambda x: abs((x + np.pi) % (2 * np.pi) - np.pi)
    for t in range(30):
        for k in range(D + 1):
            want = th0[0] + om[0] * max(t - k, 0) * DT
            assert circ(eng.get_history(k)[0] - want) < 1e-5, (t, k)
        th1 = float(eng.get_theta()[1])
        eng.step(DT)
        src = th0[0] + om[0] * max(t - d, 0) * DT
        want1 = th1 + DT * g * w * np.sin(src - th1)
        assert circ(eng.get_theta()[1] - want1) < 2e-6, t
        assert eng.get_weights()[0] == np.float32(w)


def test_set_theta_resets_history(ws1k):
    say("set_theta history reset")
    eng = make(ws1k)
    eng.step_n(DT, 20)
    new = np.linspace(0, 6, ws1k.N).astype(np.float32)
    eng.set_theta(new)
    for k in (0, 1, 25, 50):
        assert np.array_equal(eng.get_history(k), new)
    with pytest.raises(RuntimeError):
        eng.set_theta(new[:-1])


def test_nan_guards_match_oracle(ws1k):
    """NaN weight and NaN phase injected from the host: the GPU's guards give the oracle's result."""
    say("NaN guards")
    eng = make(ws1k)
    eng.step_n(DT, 5)
    w = eng.get_weights()
    w[123] = np.nan
    eng.reset_weights(w)
    th = eng.get_theta()
    th[777] = np.nan
    eng.set_theta(th)
    o = oracle_from_engine(eng, ws1k)
    o.step(DT)
    eng.step(DT)
    got_th, got_w = eng.get_theta(), eng.get_weights()
    assert np.array_equal(np.isnan(got_w), np.isnan(o.w)) and np.isfinite(got_w).all()
    assert got_w[123] == 0.0
    zero_ref = o.theta == 0.0
    assert np.array_equal(got_th == 0.0, zero_ref)
    assert zero_ref[777] and zero_ref.sum() > 1
    ok = ~zero_ref
    assert theta_rel_err(got_th[ok], o.theta[ok]) < TOL
    thh, omh = eng.get_predictor()
    assert np.isnan(omh[777]) and np.isfinite(thh[777])


def test_gain_controller_closed_form():
    """All phases equal and no coupling: r = 1 each step and the FP64 device controller
    follows g_t = g* + (g0 - g*)(1 - mu dt)^t with g* = g_rest + beta (r0 - 1)/mu."""
    from leviathan_v1_amd import LeviathanEngine

    say("gain controller")
    N = 300
    rp = np.zeros(N + 1, np.int32)
    th = np.full(N, 1.0, np.float32)
    eng = LeviathanEngine(N, 5, rp, np.zeros(0, np.int32), np.zeros(0, np.uint8), np.zeros(0, np.float32), th,
                          th.copy(), np.zeros(N, np.float32))
    r = eng.step_n(DT, 400)
    assert np.abs(r - 1.0).max() < 1e-6
    gstar = 1.5 + 0.01 * (0.5 - 1.0) / 0.01
    want = gstar + (1.5 - gstar) * (1 - 0.01 * DT) ** 400
    assert abs(eng.get_gain() - want) < 1e-6
    eng.set_gain_controller(False)
    eng.set_gain(3.25)
    eng.step_n(DT, 10)
    assert eng.get_gain() == np.float32(3.25)


def test_long_rows_barabasi_albert():
    """Heavy-tailed BA graph with a small stream capacity, so the heaviest rows take the
    whole-workgroup path: trajectory and one-step consistency against the oracle."""
    from leviathan_v1_amd import workloads
    from oracle import COracle

    wl = workloads.barabasi_albert_workload(60000, 10, 50, seed=2)
    assert wl.graph.degrees().max() > 1024
    say("BA long rows", wl.N, wl.E, wl.graph.degrees().max())
    eng = make(wl, edges_per_thread=4)
    o = COracle(*wl.engine_args(), dtype=np.float64)
    for t in range(5):
        o.step(DT)
        eng.step(DT)
        assert theta_rel_err(eng.get_theta(), o.theta) < TOL, t
        assert w_rel_err(eng.get_weights(), o.w) < TOL, t
    eng.step_n(DT, 100)
    o2 = oracle_from_engine(eng, wl)
    o2.step(DT)
    eng.step(DT)
    assert theta_rel_err(eng.get_theta(), o2.theta) < TOL
    assert w_rel_err(eng.get_weights(), o2.w) < TOL


def test_drive_matches_oracle(ws1k):
    """omega_i(t) = omega_i + b_i u_t: 10 driven steps against the FP64 oracle."""
    from oracle import COracle

    say("input drive")
    rng = np.random.default_rng(5)
    b = rng.standard_normal(ws1k.N).astype(np.float32)
    u = rng.standard_normal(10).astype(np.float32)
    eng = make(ws1k)
    eng.set_drive(b)
    eng.step_n_driven(DT, u)
    o = COracle(*ws1k.engine_args(), dtype=np.float64)
    o.drive_b = b.astype(np.float64)
    o.run(DT, 10, u_seq=u.astype(np.float64))
    assert theta_rel_err(eng.get_theta(), o.theta) < TOL
    eng.set_drive(None)
    eng.step(DT)


def test_ensemble_equals_separate_engines():
    """Four networks of uneven size in one launch == four engines run one by one (bitwise)."""
    from leviathan_v1_amd import LeviathanEngine, graph, workloads

    say("ensemble vs separate")
    sizes = [500, 1200, 3000, 777]
    wls = [workloads.watts_strogatz_workload(n, 20, 0.1, 50, seed=10 + i) for i, n in enumerate(sizes)]
    g, offs = graph.disjoint_union([w.graph for w in wls])
    cat = lambda f: np.concatenate([getattr(w, f) for w in wls])
    g0 = [0.5, 1.0, 2.5, 4.0]
    ctl = [False, True, False, True]
    ens = LeviathanEngine(g.num_nodes, 50, g.row_ptr, g.col_idx, cat("delays"), cat("weights"), cat("theta"),
                          cat("theta_hat"), cat("omega"), n_nets=4, net_offsets=offs, g0=g0, controller=ctl)
    r_ens = ens.step_n(DT, 120)
    th = ens.get_theta()
    for k, w in enumerate(wls):
        e = LeviathanEngine(*w.engine_args(), params=dict(g0=g0[k], gain_controller=ctl[k]))
        r = e.step_n(DT, 120)
        assert np.array_equal(r, r_ens[:, k]), k
        assert np.array_equal(e.get_theta(), th[offs[k]:offs[k + 1]]), k
        assert e.get_gain() == ens.get_gain_net(k)


def test_observatory_surface():
    from leviathan_v1_amd import LeviathanObservatory

    say("observatory")
    obs = LeviathanObservatory(N=3000, k=20, max_delay=50, seed=1)
    obs.run_baseline(num_steps=50)
    r = obs.run_with_stimulus(num_steps=40, stimulus_start=10)
    assert len(r) == 40 and obs.step_count == 90
    obs.step_n(100)
    st = obs.statistics()
    assert st["steps"] == 190 and 0 <= st["min_r"] <= st["max_r"] <= 1
    snap = obs.get_phase_snapshot()
    assert snap.dtype == np.float32 and snap.shape == (3000,)
    obs.set_phase_snapshot(snap.astype(np.float64))
    assert "TOTAL (Measured)" in obs.get_vram_report()
    assert obs.get_vram_usage() >= obs.nnz * 9 + 3000 * (52 * 4 + 16)


# ---------------------------------------------------------------------------
# edge cases of the row blocks: no edges at all, rows without incoming edges,
# a hub beyond one block's capacity next to isolated rows
# ---------------------------------------------------------------------------

def _ragged_workload(n, src, dst, seed):
    from leviathan_v1_amd import graph, workloads

    g = graph.from_edges(n, np.asarray(src, np.int32), np.asarray(dst, np.int32))
    delays, weights, theta, thh, omega = workloads.init_state(g, 50, seed)
    return workloads.Workload("ragged", g, 50, delays, weights, theta, thh, omega)


@pytest.mark.parametrize("schedule,ept", [("direct", 4), ("direct", 16), ("binned", 8), ("staged", 4),
                                          ("staged", 16)])
def test_edgeless_and_ragged_graphs(schedule, ept):
    """E = 0 (phases advance by omega alone), and a graph whose rows 1200.. have no
    incoming edges while node 0 is joined to 1..1500 (a row longer than a 4-edge-per-lane
    block): one step against the FP64 oracle from the engine's state, in every schedule."""
    say("ragged", schedule, ept)
    rng = np.random.default_rng(5)
    a = rng.integers(1, 1200, 3000)
    b = rng.integers(1, 1200, 3000)
    keep = a != b
    cases = [_ragged_workload(500, [], [], 1),
             _ragged_workload(2000, np.concatenate([np.zeros(1500, int), a[keep]]),
                              np.concatenate([np.arange(1, 1501), b[keep]]), 2)]
    for wl in cases:
        eng = make(wl, schedule=schedule, edges_per_thread=ept, history="u16")
        r = eng.step_n(DT, 7)
        assert np.isfinite(r).all()
        o = oracle_from_engine(eng, wl)
        o.step(DT)
        eng.step(DT)
        assert theta_rel_err(eng.get_theta(), o.theta) < TOL, (wl.E, schedule)
        if wl.E:
            assert w_rel_err(eng.get_weights(), o.w) < TOL
        eng.close()


@pytest.mark.parametrize("history", ["u16", "f32"])
def test_staged_regroup_many_blocks_per_bucket(history):
    """A sparse graph (1.2M nodes, 600k random undirected edges): about 2,000 row blocks
    per 512k-edge bucket, so regroup chunks are cut at the run limit; one step of the
    staged schedule against the FP64 oracle, and the same step as the direct schedule."""
    say("staged regroup, sparse", history)
    rng = np.random.default_rng(9)
    n = 1_200_000
    a = rng.integers(0, n, 600_000)
    b = rng.integers(0, n, 600_000)
    keep = a != b
    wl = _ragged_workload(n, a[keep], b[keep], 3)
    eng = make(wl, schedule="staged", history=history)
    ref = make(wl, history=history)
    eng.step_n(DT, 5)
    ref.step_n(DT, 5)
    assert np.array_equal(eng.get_weights(), ref.get_weights())
    o = oracle_from_engine(eng, wl)
    o.step(DT)
    eng.step(DT)
    assert theta_rel_err(eng.get_theta(), o.theta) < TOL
    assert w_rel_err(eng.get_weights(), o.w) < TOL
    eng.close()
    ref.close()


@pytest.mark.parametrize("history,schedule", [("f32", "direct"), ("u16", "staged"), ("u16", "binned")])
def test_checkpoint_restore_continues_exactly(ws1k, tmp_path, history, schedule):
    """A fresh engine restored from a checkpoint file (theta, predictor, weights, every
    history slot, FP64 gain, step count) continues bit for bit like the engine it came from."""
    from leviathan_v1_amd import checkpoint

    say("checkpoint", history, schedule)
    a = make(ws1k, history=history, schedule=schedule)
    a.step_n(DT, 25)
    p = tmp_path / "state.npz"
    g = ws1k.graph
    checkpoint.save(p, a.checkpoint(), g.row_ptr, g.col_idx, ws1k.delays)
    ra = a.step_n(DT, 30)
    b = make(ws1k, history=history, schedule=schedule)
    b.step_n(DT, 3)  # any prior state is overwritten
    b.restore(checkpoint.load(p, g.row_ptr, g.col_idx, ws1k.delays))
    assert b.step_count == 25
    rb = b.step_n(DT, 30)
    assert np.array_equal(ra, rb)
    assert np.array_equal(a.get_theta(), b.get_theta())
    assert np.array_equal(a.get_weights(), b.get_weights())
    assert np.array_equal(a.get_gains(), b.get_gains())


def test_checkpoint_restore_ensemble():
    """Ensembles: per-network FP64 gains (controller on for one network) survive the restore."""
    from leviathan_v1_amd import LeviathanEngine, graph, workloads

    say("checkpoint ensemble")
    wls = [workloads.watts_strogatz_workload(n, 20, 0.1, 50, seed=60 + i) for i, n in enumerate([700, 1500])]
    g, offs = graph.disjoint_union([w.graph for w in wls])
    cat = lambda f: np.concatenate([getattr(w, f) for w in wls])
    args = (g.num_nodes, 50, g.row_ptr, g.col_idx, cat("delays"), cat("weights"), cat("theta"),
            cat("theta_hat"), cat("omega"))
    kw = dict(n_nets=2, net_offsets=offs, g0=[0.7, 2.5], controller=[True, False], history="u16")
    a = LeviathanEngine(*args, **kw)
    a.step_n(DT, 40)
    st = a.checkpoint()
    ra = a.step_n(DT, 20)
    b = LeviathanEngine(*args, **kw)
    b.restore(st)
    rb = b.step_n(DT, 20)
    assert np.array_equal(ra, rb)
    assert np.array_equal(a.get_gains(), b.get_gains())
    assert a.get_gains()[0] != 0.7  # the controller moved the first network's gain


# ---------------------------------------------------------------------------
# binned (two-phase) schedule
# ---------------------------------------------------------------------------

@pytest.fixture(scope="module")
def ba200k():
    from leviathan_v1_amd import workloads

    wl = workloads.barabasi_albert_workload(200_000, 10, 50, seed=4)
    assert wl.graph.degrees().max() > 2048  # exercises single-row buckets and long sources
    return wl


@pytest.mark.parametrize("schedule", ["binned", "staged"])
@pytest.mark.parametrize("history", ["f32", "u16"])
@pytest.mark.parametrize("layout", ["node", "slot"])
def test_binned_equals_direct(ba200k, history, layout, schedule):
    """The two-phase schedules share the direct step's per-row arithmetic and summation
    order: phases and weights agree to FP32 rounding after 40 steps (the graph has rows
    and sources beyond one block's capacity)."""
    say(schedule, "vs direct", history, layout)
    d = make(ba200k, history=history, layout=layout)
    b = make(ba200k, history=history, layout=layout, schedule=schedule)
    assert b.get_schedule() == (schedule, 8)
    rd = d.step_n(DT, 40)
    rb = b.step_n(DT, 40)
    assert np.abs(rd - rb).max() < 1e-6
    assert theta_rel_err(b.get_theta(), d.get_theta()) < 1e-6
    assert np.abs(b.get_weights() - d.get_weights()).max() < 1e-6


@pytest.fixture(scope="module")
def ws3k_d0():
    """WS N=3,000 with max_delay 5 and delays uniform in [0, 5]: delay 0 reads the phases the
    previous step wrote, delay 5 the oldest slot; the rows hold 20 edges, each source's edges
    span all six delays."""
    import dataclasses

    from leviathan_v1_amd import workloads

    wl = workloads.watts_strogatz_workload(3000, 20, 0.1, 5, seed=21)
    rng = np.random.default_rng(5)
    return dataclasses.replace(wl, delays=rng.integers(0, 6, wl.E).astype(np.uint8))


@pytest.mark.parametrize("schedule", ["direct", "binned", "staged"])
@pytest.mark.parametrize("layout", ["node", "slot"])
@pytest.mark.parametrize("history", ["f32", "u16"])
def test_delay_zero_every_schedule_vs_fp64_oracle(ws3k_d0, schedule, layout, history):
    """Every schedule and ring layout with delays 0..D against one FP64 oracle step taken from
    the engine's state after 12 steps (more than two trips round the 7-slot ring)."""
    say("delay 0..5", schedule, layout, history)
    eng = make(ws3k_d0, schedule=schedule, layout=layout, history=history)
    eng.step_n(DT, 12)
    o = oracle_from_engine(eng, ws3k_d0)
    o.step(DT)
    eng.step(DT)
    assert theta_rel_err(eng.get_theta(), o.theta) < TOL
    assert w_rel_err(eng.get_weights(), o.w) < TOL


@pytest.mark.parametrize("schedule,bucket_edges,ept", [("binned", 0, 8), ("binned", 2048, 8), ("staged", 0, 4),
                                                       ("staged", 0, 16)])
def test_binned_one_step_consistency(ws1k, schedule, bucket_edges, ept):
    """Two-phase schedules (default and minimum-size buckets; staged with the smallest and
    largest row blocks) against the FP64 oracle."""
    say(schedule, "one-step", bucket_edges, ept)
    eng = make(ws1k, schedule=schedule, bucket_edges=bucket_edges, history="u16", layout="slot",
               edges_per_thread=ept)
    eng.step_n(DT, 33)
    o = oracle_from_engine(eng, ws1k)
    o.step(DT)
    eng.step(DT)
    assert theta_rel_err(eng.get_theta(), o.theta) < TOL
    assert w_rel_err(eng.get_weights(), o.w) < TOL


@pytest.mark.parametrize("schedule", ["binned", "staged"])
def test_binned_trajectory_vs_golden(ws1k, golden1k, schedule):
    say(schedule, "trajectory")
    eng = make(ws1k, schedule=schedule)
    eng.step_n(DT, 10)
    assert theta_rel_err(eng.get_theta(), golden1k["theta_10"]) < TOL
    assert w_rel_err(eng.get_weights(), golden1k["w_10"]) < TOL


def test_autotune_restores_state_and_switches(ba200k):
    """autotune() times every schedule from the current state, restores it, keeps the fastest."""
    say("autotune")
    a = make(ba200k, schedule="auto")
    ref = make(ba200k)
    a.step_n(DT, 5)
    ref.step_n(DT, 5)
    times = a.autotune(DT, steps=3)
    say("autotune ms/step", times)
    assert set(times) == {f"{k}/{e}" for k in ("direct", "binned", "staged") for e in (4, 8, 16)}
    assert a.step_count == 5
    assert np.array_equal(a.get_theta(), ref.get_theta())
    assert np.array_equal(a.get_weights(), ref.get_weights())
    best = min(times, key=times.get)
    assert a.schedule_name == best
    ra = a.step_n(DT, 20)
    rr = ref.step_n(DT, 20)
    assert np.abs(ra - rr).max() < 1e-6
    assert theta_rel_err(a.get_theta(), ref.get_theta()) < 1e-6
    for s, e in [("direct", 4), ("binned", 16), ("staged", 8), ("direct", 16), ("staged", 16), ("binned", 4),
                 ("staged", 4)]:
        a.set_schedule(s, e)
        a.step_n(DT, 3)
        ref.step_n(DT, 3)
        assert np.array_equal(a.get_weights(), ref.get_weights()), (s, e)
    assert theta_rel_err(a.get_theta(), ref.get_theta()) < 1e-6
    w = ref.get_weights()
    w[::7] = 0.25
    a.reset_weights(w)
    assert np.array_equal(a.get_weights(), w)
    with pytest.raises(RuntimeError):
        ref.set_schedule("binned")


@pytest.mark.parametrize("schedule", ["binned", "staged"])
def test_binned_ensemble_equals_separate(schedule):
    """Ensemble in a two-phase schedule == separate direct engines."""
    from leviathan_v1_amd import LeviathanEngine, graph, workloads

    say(schedule, "ensemble")
    sizes = [900, 2500, 1300]
    wls = [workloads.watts_strogatz_workload(n, 20, 0.1, 50, seed=30 + i) for i, n in enumerate(sizes)]
    g, offs = graph.disjoint_union([w.graph for w in wls])
    cat = lambda f: np.concatenate([getattr(w, f) for w in wls])
    g0 = [0.3, 2.0, 3.5]
    ens = LeviathanEngine(g.num_nodes, 50, g.row_ptr, g.col_idx, cat("delays"), cat("weights"), cat("theta"),
                          cat("theta_hat"), cat("omega"), n_nets=3, net_offsets=offs, g0=g0,
                          controller=[False] * 3, schedule=schedule, bucket_edges=4096)
    r_ens = ens.step_n(DT, 60)
    th = ens.get_theta()
    for k, w in enumerate(wls):
        e = LeviathanEngine(*w.engine_args(), params=dict(g0=g0[k], gain_controller=False))
        r = e.step_n(DT, 60)
        assert np.abs(r - r_ens[:, k]).max() < 1e-6, k
        assert theta_rel_err(e.get_theta(), th[offs[k]:offs[k + 1]]) < 1e-6, k


@pytest.mark.parametrize("history,schedule", [("f32", "direct"), ("u16", "direct"), ("u16", "binned"),
                                              ("u16", "staged")])
def test_mid_size_long_run_r_statistics(history, schedule):
    """WS N=100,000 k=20 p=0.1 with plasticity and homeostasis, 10,000 steps: windowed
    (500-step) mean r and r variance within 5e-3 of the FP64 oracle's run (golden fixture)."""
    from leviathan_v1_amd import workloads
    from tests.golden.make_golden import digest

    say("ws100k 10k steps", history, schedule)
    gold = np.load(os.path.join(HERE, "golden", "ws100k_fp64.npz"))
    wl = workloads.mid_plain(0)
    assert str(gold["digest"]) == digest(wl), "workload generator changed"
    eng = make(wl, history=history, schedule=schedule)
    r = eng.step_n(DT, 10_000)
    m, v = windowed(r, 1000, 500)
    mr, vr = windowed(gold["r"], 1000, 500)
    say("max |mean r diff|", float(np.abs(m - mr).max()), "max |var diff|", float(np.abs(v - vr).max()),
        "g", eng.get_gain(), float(gold["g_final"]))
    assert np.abs(m - mr).max() < 5e-3
    assert np.abs(v - vr).max() < 5e-3


def test_ensemble_r_of_K_vs_oracle():
    """Three networks (K = 0.5, 2, 4, controller off) in one batched engine: per-network
    windowed r statistics over 2,000 steps within 5e-3 of each network's FP64 oracle run."""
    from leviathan_v1_amd import workloads
    from leviathan_v1_amd.ensemble import LeviathanEnsemble
    from oracle import COracle, Params

    say("ensemble r(K) vs oracle")
    gains = [0.5, 2.0, 4.0]
    wls = [workloads.watts_strogatz_workload(2000, 20, 0.1, 50, seed=50 + i) for i in range(3)]
    ens = LeviathanEnsemble(wls, gains, history="u16", schedule="auto")
    ens.engine.autotune(DT, steps=2)
    K, r_mean, r_std = ens.phase_sweep(0, 2000)
    ens2 = LeviathanEnsemble(wls, gains)
    r_all = ens2.step_n(2000)
    for k, w in enumerate(wls):
        o = COracle(*w.engine_args(), params=Params(g0=gains[k], gain_controller=False), dtype=np.float64)
        r_ref = o.run(DT, 2000)
        m, v = windowed(r_all[:, k], 500, 500)
        mr, vr = windowed(r_ref, 500, 500)
        assert np.abs(m - mr).max() < 5e-3 and np.abs(v - vr).max() < 5e-3, k
        assert abs(r_mean[k] - r_ref.mean()) < 5e-3, k
