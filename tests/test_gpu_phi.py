"""Hub transition statistics on the device (include/leviathan_phi.h) against the oracle.

* counts, row totals, last state and call count: bit-exact (integer work);
* state-by-node TPM: bit-exact (one IEEE FP64 division of exact integer sums);
* phi per bipartition: within 1e-13 of the FP64 oracle, both ~0 (the reference's formula
  is zero in exact arithmetic: p_A p_B multiplies the factors of p_int), exactly 0 for
  the uniform TPM; H = 1 gives 1e30 and mask 0 (no bipartition);
* attached to an engine, every step records the state of the new phases on the device
  (graph-captured chunks and single steps alike) without changing the step: the
  phases match an engine without the handle bit for bit, and the counts match the oracle
  fed with the phases the reference's host round trip would copy (get_theta per step).
"""
import ctypes

import numpy as np
import pytest

from gpu_helpers import say
from oracle import phi as P

pytestmark = pytest.mark.gpu

DT = 0.05


def _wl(n=2000, seed=7):
    from leviathan_v1_amd import workloads

    return workloads.watts_strogatz_workload(n, 20, 0.1, 50, seed=seed)


def _engine(wl, **kw):
    from leviathan_v1_amd import LeviathanEngine

    return LeviathanEngine(*wl.engine_args(), **kw)


def _worker(hubs):
    from leviathan_v1_amd.phi import GPUPhiWorker

    return GPUPhiWorker(hubs)


def _check_stats(w, o):
    c, t = w.counts()
    s, n = w.state()
    assert np.array_equal(c, o.counts)
    assert np.array_equal(t, o.row_totals)
    assert (s, n) == (o.prev, o.calls)
    assert w.get_transition_count() == o.calls


def _check_compute(w, o):
    ex = w.compute_ex()
    sbn = o.sbn_tpm()
    assert np.array_equal(ex["sbn_tpm"], sbn)
    part = P.phi_partitions(sbn, o.H)
    assert ex["partitions"].shape == part.shape
    if part.size:
        assert np.abs(ex["partitions"] - part).max() <= 1e-13
        assert ex["partitions"].max() <= 1e-13
        k = ex["mip_mask"] - 1
        assert ex["phi"] == ex["partitions"][k] == ex["partitions"].min()
        assert k == int(np.argmin(ex["partitions"]))
    else:
        assert (ex["phi"], ex["mip_mask"]) == (P.INVALID, 0)
    assert w.compute() == np.float32(ex["phi"])
    return ex


@pytest.mark.parametrize("hubs", [[0], [5, 1], [3, 0, 7, 5], list(range(10))])
def test_host_accumulate_matches_oracle(hubs):
    say("phi host accumulate hubs", hubs)
    rng = np.random.default_rng(len(hubs))
    w = _worker(hubs)
    o = P.PhiOracle(hubs)
    T = 600
    for t in range(T):
        th = rng.uniform(-7, 14, 12).astype(np.float32)
        if t % 50 == 0:
            th[hubs[0]] = np.float32(np.pi)
        w.accumulate(th)
        o.accumulate(th)
    _check_stats(w, o)
    _check_compute(w, o)
    w.reset()
    o.reset()
    _check_stats(w, o)
    ex = _check_compute(w, o)  # uniform TPM: every bipartition exactly 0
    assert np.all(ex["partitions"] == 0.0)
    w.close()


def test_device_accumulate_ten_hubs_many_transitions():
    """S = 1024 states, 20,000 phase vectors already in device memory (the reference's
    phi_accumulate(d_theta) form), then the full compute against the oracle."""
    say("phi device accumulate 10 hubs")
    hip = ctypes.CDLL("libamdhip64.so")
    rng = np.random.default_rng(11)
    T, H = 20_000, 10
    # phases biased per hub so that the table is far from uniform
    bias = rng.uniform(0.2, 0.8, H)
    on = rng.random((T, H)) < bias
    th = np.where(on, rng.uniform(3.2, 6.2, (T, H)), rng.uniform(0.0, 3.1, (T, H))).astype(np.float32)
    ptr = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(th.nbytes)) == 0
    try:
        assert hip.hipMemcpy(ptr, ctypes.c_void_p(th.ctypes.data), ctypes.c_size_t(th.nbytes), 1) == 0
        w = _worker(list(range(H)))
        o = P.PhiOracle(list(range(H)))
        for t in range(T):
            w.accumulate_device(ptr.value + 4 * H * t)
            o.accumulate(th[t])
        _check_stats(w, o)
        _check_compute(w, o)
        w.close()
    finally:
        hip.hipFree(ptr)


@pytest.mark.parametrize("history,schedule,K", [("f32", "direct", 1), ("u16", "staged", 1), ("u16", "staged", 4)])
def test_attached_engine_records_every_step(history, schedule, K):
    """The engine's step_n (paired row passes for the direct schedule and for four steps per scatter
    pass: each pair's two finalizes record the phases of their own step) against single steps with
    the reference's host round trip."""
    from leviathan_v1_amd.phi import top_degree_hubs

    say("phi attached", history, schedule, K)
    wl = _wl()
    hubs = top_degree_hubs(wl.graph.row_ptr, wl.graph.col_idx, 6)
    a = _engine(wl, history=history, schedule=schedule)
    b = _engine(wl, history=history, schedule=schedule)
    if K > 1:
        a.set_scatter_steps(K)
    assert a.row_steps == 2
    w = _worker(hubs)
    w.attach(a)
    o = P.PhiOracle(hubs)
    ra = np.concatenate([a.step_n(DT, 250), [a.step(DT)], a.step_n(DT, 30)])
    rb = []
    for _ in range(281):  # the reference's host round trip: one snapshot per step
        rb.append(b.step(DT))
        o.accumulate(b.get_theta())
    assert np.array_equal(ra, np.float32(rb))
    assert np.array_equal(a.get_theta(), b.get_theta())
    _check_stats(w, o)
    _check_compute(w, o)
    # host accumulation is refused while the engine records
    with pytest.raises(RuntimeError, match="attached"):
        w.accumulate(a.get_theta())
    # autotune's trial steps are rolled back and record nothing
    a.autotune(DT, steps=2)
    _check_stats(w, o)
    # after detach the engine steps without recording
    w.detach()
    a.step_n(DT, 20)
    _check_stats(w, o)
    w.close()
    a.close()
    b.close()


def test_attached_ensemble_engine():
    """An ensemble of 40 networks (its finalize is a grid of workgroups): hubs in several networks
    record exactly what the oracle records from the phases after every step."""
    from leviathan_v1_amd import LeviathanEngine, graph, workloads

    say("phi attached ensemble")
    wls = [workloads.watts_strogatz_workload(300 + 20 * k, 20, 0.1, 50, seed=40 + k) for k in range(40)]
    g, offs = graph.disjoint_union([w.graph for w in wls])
    cat = lambda f: np.concatenate([getattr(w, f) for w in wls])
    args = (g.num_nodes, 50, g.row_ptr, g.col_idx, cat("delays"), cat("weights"), cat("theta"), cat("theta_hat"),
            cat("omega"))
    kw = dict(n_nets=40, net_offsets=offs, g0=list(np.linspace(0, 4, 40)), controller=[False] * 40)
    a = LeviathanEngine(*args, **kw)
    b = LeviathanEngine(*args, **kw)
    hubs = [int(offs[k]) + 7 for k in (0, 9, 17, 25, 33, 39)]
    w = _worker(hubs)
    w.attach(a)
    o = P.PhiOracle(hubs)
    a.step_n(DT, 150)
    for _ in range(150):
        b.step(DT)
        o.accumulate(b.get_theta())
    assert np.array_equal(a.get_theta(), b.get_theta())
    _check_stats(w, o)
    w.close()
    a.close()
    b.close()


def test_attach_rules_and_free_while_attached():
    say("phi attach rules")
    wl = _wl(500, seed=3)
    a = _engine(wl)
    b = _engine(wl)
    w = _worker([1, 2, 3])
    w.attach(a)
    w2 = _worker([4, 5])
    with pytest.raises(RuntimeError, match="another engine"):
        w.attach(b)
    w2.attach(a)  # replaces w on engine a
    a.step_n(DT, 10)
    assert w.get_transition_count() == 0 and w2.get_transition_count() == 10
    too_far = _worker([wl.N])
    with pytest.raises(RuntimeError, match="out of range"):
        too_far.attach(b)
    w2.close()  # frees while attached: the engine detaches it
    a.step_n(DT, 10)
    w.accumulate(a.get_theta())  # w was detached by the replacement
    assert w.get_transition_count() == 1
    a.close()
    b.close()
    w.close()
    too_far.close()


def test_checkpoint_restore_continues_identically():
    say("phi checkpoint")
    wl = _wl(1000, seed=5)
    hubs = [10, 20, 30, 40, 50]
    a = _engine(wl)
    w = _worker(hubs)
    w.attach(a)
    a.step_n(DT, 120)
    ck_e = a.checkpoint()
    ck_w = w.checkpoint()
    a.step_n(DT, 80)
    want_c, want_t = w.counts()
    want_s = w.state()
    b = _engine(wl)
    b.restore(ck_e)
    w2 = _worker(hubs)
    w2.restore(ck_w)
    w2.attach(b)
    b.step_n(DT, 80)
    c, t = w2.counts()
    assert np.array_equal(c, want_c) and np.array_equal(t, want_t) and w2.state() == want_s
    with pytest.raises(RuntimeError, match="sums to"):
        bad = dict(ck_w, counts=ck_w["counts"] + np.eye(32, dtype=np.int64))
        w2.restore(bad)
    for x in (w, w2, a, b):
        x.close()
