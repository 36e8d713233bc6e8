"""Topology builders: structure, reproducibility, and agreement with networkx
where the algorithm is deterministic (p = 0 ring lattice)."""
import networkx as nx
import numpy as np
import pytest

from leviathan_v1_amd import graph, workloads


def _edges(g):
    rows = np.repeat(np.arange(g.num_nodes), g.degrees())
    return set(zip(rows.tolist(), g.col_idx.tolist()))


@pytest.mark.parametrize("n,k", [(20, 4), (101, 10), (1000, 20)])
def test_ws_p0_is_networkx_lattice(n, k):
    g = graph.watts_strogatz(n, k, 0.0, seed=1)
    ref = nx.to_scipy_sparse_array(nx.watts_strogatz_graph(n, k, 0.0), format="csr")
    ref.sort_indices()
    assert np.array_equal(g.row_ptr, ref.indptr)
    assert np.array_equal(g.col_idx, ref.indices)


@pytest.mark.parametrize("p", [0.1, 0.5, 1.0])
def test_ws_rewiring_keeps_simple_symmetric_graph(p):
    n, k = 2000, 20
    g = graph.watts_strogatz(n, k, p, seed=7)
    assert g.num_edges == n * k  # rewiring moves edges, never adds or drops them
    e = _edges(g)
    assert all((j, i) in e for (i, j) in e)  # both directions stored
    assert all(i != j for (i, j) in e)  # no self loops
    for i in range(0, n, 97):
        row = g.col_idx[g.row_ptr[i]:g.row_ptr[i + 1]]
        assert np.all(np.diff(row) > 0)  # sorted, no duplicates
    # fraction of non-lattice edges ~ p (each lattice edge rewired with probability p)
    lattice = {(i, (i + d) % n) for i in range(n) for d in range(1, k // 2 + 1)}
    lattice |= {(j, i) for (i, j) in lattice}
    frac = len(e - lattice) / len(e)
    assert abs(frac - p * (1 - 2 * k / (2 * n))) < 0.05


def test_ws_seeded_and_distinct():
    a = graph.watts_strogatz(5000, 20, 0.1, seed=3)
    b = graph.watts_strogatz(5000, 20, 0.1, seed=3)
    c = graph.watts_strogatz(5000, 20, 0.1, seed=4)
    assert np.array_equal(a.col_idx, b.col_idx)
    assert not np.array_equal(a.col_idx, c.col_idx)


@pytest.mark.parametrize("n,m", [(50, 3), (5000, 10)])
def test_ba_structure(n, m):
    g = graph.barabasi_albert(n, m, seed=11)
    assert g.num_edges == 2 * (m + m * (n - m - 1))
    deg = g.degrees()
    assert deg[m + 1:].min() >= m  # every added node brings m distinct edges
    e = _edges(g)
    assert all((j, i) in e for (i, j) in e)
    assert all(i != j for (i, j) in e)
    for i in range(n):
        row = g.col_idx[g.row_ptr[i]:g.row_ptr[i + 1]]
        assert np.all(np.diff(row) > 0)


def test_ba_heavy_tail_matches_networkx_statistics():
    """Degree distribution of the native builder vs networkx's generator (same algorithm,
    different random stream): the tails agree statistically."""
    n, m = 20000, 10
    g = graph.barabasi_albert(n, m, seed=5)
    ref = nx.barabasi_albert_graph(n, m, seed=5)
    d_ref = np.array([d for _, d in ref.degree()])
    d = g.degrees()
    assert d.sum() == d_ref.sum()
    for q in (0.5, 0.9, 0.99):
        assert abs(np.quantile(d, q) - np.quantile(d_ref, q)) <= max(2, 0.15 * np.quantile(d_ref, q))
    assert 0.4 < d.max() / d_ref.max() < 2.5


def test_disjoint_union_offsets():
    gs = [graph.watts_strogatz(n, 4, 0.2, seed=n) for n in (10, 30, 7)]
    u, offs = graph.disjoint_union(gs)
    assert offs.tolist() == [0, 10, 40, 47]
    assert u.num_edges == sum(g.num_edges for g in gs)
    for i, g in enumerate(gs):
        lo, hi = offs[i], offs[i + 1]
        sub = u.col_idx[u.row_ptr[lo]:u.row_ptr[hi]]
        assert sub.min() >= lo and sub.max() < hi


def test_workload_distributions():
    wl = workloads.small_plain(0)
    assert wl.delays.min() >= 1 and wl.delays.max() == 50
    assert wl.weights.dtype == np.float32 and 0 <= wl.weights.min() and wl.weights.max() < 1
    assert 0 <= wl.theta.min() and wl.theta.max() < 2 * np.pi
    assert abs(wl.omega.mean()) < 0.15 and abs(wl.omega.std() - 1) < 0.1
    assert np.array_equal(wl.theta, wl.theta_hat)


# ---------------------------------------------------------------------------
# relabelling and graph files
# ---------------------------------------------------------------------------

def test_permute_maps_every_edge_and_inverts():
    from leviathan_v1_amd import graph

    g = graph.barabasi_albert(3000, 4, seed=11)
    rng = np.random.default_rng(0)
    perm = rng.permutation(g.num_nodes).astype(np.int32)
    h, emap = graph.permute(g, perm)
    # new entry (row i, column c) is old entry (row perm[i], column perm[c])
    rows = np.repeat(np.arange(h.num_nodes), np.diff(h.row_ptr))
    old_rows = np.repeat(np.arange(g.num_nodes), np.diff(g.row_ptr))
    assert np.array_equal(old_rows[emap], perm[rows])
    assert np.array_equal(g.col_idx[emap], perm[h.col_idx])
    assert np.array_equal(np.sort(emap), np.arange(g.num_edges))
    # columns ascending within rows
    d = np.diff(h.col_idx)
    assert (d[np.diff(rows) == 0] > 0).all()
    inv = np.empty_like(perm)
    inv[perm] = np.arange(perm.size, dtype=np.int32)
    back, emap2 = graph.permute(h, inv)
    assert np.array_equal(back.row_ptr, g.row_ptr) and np.array_equal(back.col_idx, g.col_idx)
    assert np.array_equal(emap[emap2], np.arange(g.num_edges))
    with pytest.raises(ValueError):
        graph.permute(g, np.zeros(g.num_nodes, dtype=np.int32))


def test_orderings():
    from leviathan_v1_amd import graph

    g = graph.barabasi_albert(2000, 3, seed=2)
    o = graph.degree_order(g)
    assert (np.diff(graph.permute(g, o)[0].degrees()) <= 0).all()
    # a ring lattice under a random labelling: RCM brings the bandwidth back down to ~k
    ring = graph.watts_strogatz(4000, 6, 0.0, seed=0)
    rng = np.random.default_rng(1)
    shuffled, _ = graph.permute(ring, rng.permutation(4000).astype(np.int32))
    r, _ = graph.permute(shuffled, graph.rcm_order(shuffled))

    def bandwidth(h):
        rows = np.repeat(np.arange(h.num_nodes), np.diff(h.row_ptr))
        return int(np.abs(rows - h.col_idx).max())

    assert bandwidth(shuffled) > 1000 and bandwidth(r) <= 12


def test_graph_file_round_trip(tmp_path):
    from leviathan_v1_amd import graph

    g = graph.watts_strogatz(500, 8, 0.2, seed=3)
    delays = np.arange(g.num_edges, dtype=np.uint8)
    graph.save(tmp_path / "g.npz", g, delays=delays)
    h, extra = graph.load(tmp_path / "g.npz")
    assert np.array_equal(h.row_ptr, g.row_ptr) and np.array_equal(h.col_idx, g.col_idx)
    assert np.array_equal(extra["delays"], delays)
    bad = g.col_idx.copy()
    bad[5] = 500
    np.savez(tmp_path / "bad.npz", row_ptr=g.row_ptr, col_idx=bad)
    with pytest.raises(ValueError):
        graph.load(tmp_path / "bad.npz")


def test_relabelled_workload_steps_like_the_original():
    """FP64 oracle: one step of a relabelled workload equals the original's step, relabelled
    (row sums run in a different order, so equality is to rounding)."""
    import oracle
    from leviathan_v1_amd import graph, workloads

    wl = workloads.barabasi_albert_workload(1500, 4, 20, seed=5)
    perm = graph.degree_order(wl.graph)
    rl = workloads.relabel(wl, perm)
    a = oracle.COracle(*wl.engine_args(), dtype=np.float64)
    b = oracle.COracle(*rl.engine_args(), dtype=np.float64)
    for _ in range(5):
        a.step(0.05)
        b.step(0.05)
    assert np.abs(oracle.circ_err(b.theta, a.theta[perm])).max() < 1e-12
    _, emap = graph.permute(wl.graph, perm)
    assert np.abs(b.w - a.w[emap]).max() < 1e-12
