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
