"""Seeded topology builders (native, ``libleviathan_graph.so``).

Replaces the reference's networkx construction (python/leviathan_h100.py:34-45):
``nx.watts_strogatz_graph`` + ``nx.to_scipy_sparse_array(format="csr")``.  The
generators restate networkx 3.4.2's algorithms with a seeded xoshiro256**
stream; the output is the incoming-edge CSR (``row_ptr`` int32[N+1],
``col_idx`` int32[E]) with both directions of every undirected edge and sorted
columns.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import _build

_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_build.GRAPH_SO):
            raise RuntimeError(
                f"{_build.GRAPH_SO} is missing: build it first (python -m leviathan_v1_amd._build graph)"
            )
        lib = ctypes.CDLL(_build.GRAPH_SO)
        vp = ctypes.c_void_p
        lib.lev_graph_watts_strogatz.restype = vp
        lib.lev_graph_watts_strogatz.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_double, ctypes.c_uint64]
        lib.lev_graph_barabasi_albert.restype = vp
        lib.lev_graph_barabasi_albert.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_uint64]
        lib.lev_graph_from_edges.restype = vp
        lib.lev_graph_from_edges.argtypes = [ctypes.c_int32, ctypes.c_int64, vp, vp]
        lib.lev_graph_num_nodes.restype = ctypes.c_int32
        lib.lev_graph_num_nodes.argtypes = [vp]
        lib.lev_graph_num_entries.restype = ctypes.c_int64
        lib.lev_graph_num_entries.argtypes = [vp]
        lib.lev_graph_copy_csr.restype = None
        lib.lev_graph_copy_csr.argtypes = [vp, vp, vp]
        lib.lev_graph_free.restype = None
        lib.lev_graph_free.argtypes = [vp]
        lib.lev_graph_last_error.restype = ctypes.c_char_p
        lib.lev_graph_permute.restype = ctypes.c_int
        lib.lev_graph_permute.argtypes = [ctypes.c_int32, vp, vp, vp, vp, vp, vp]
        _lib = lib
    return _lib


@dataclass
class CSRGraph:
    """Incoming-edge CSR: row i lists the sources j of edges j -> i."""

    row_ptr: np.ndarray
    col_idx: np.ndarray

    @property
    def num_nodes(self) -> int:
        return int(self.row_ptr.shape[0] - 1)

    @property
    def num_edges(self) -> int:
        return int(self.col_idx.shape[0])

    def degrees(self) -> np.ndarray:
        return np.diff(self.row_ptr)


def _take(h) -> CSRGraph:
    lib = _load()
    if not h:
        raise RuntimeError(lib.lev_graph_last_error().decode())
    try:
        n = lib.lev_graph_num_nodes(h)
        e = lib.lev_graph_num_entries(h)
        row_ptr = np.empty(n + 1, dtype=np.int32)
        col_idx = np.empty(e, dtype=np.int32)
        lib.lev_graph_copy_csr(h, row_ptr.ctypes.data, col_idx.ctypes.data)
    finally:
        lib.lev_graph_free(h)
    return CSRGraph(row_ptr, col_idx)


def watts_strogatz(n: int, k: int, p: float, seed: int) -> CSRGraph:
    """Watts-Strogatz small world (networkx watts_strogatz_graph's algorithm)."""
    return _take(_load().lev_graph_watts_strogatz(int(n), int(k), float(p), int(seed) & (2**64 - 1)))


def barabasi_albert(n: int, m: int, seed: int) -> CSRGraph:
    """Barabasi-Albert graph (networkx barabasi_albert_graph's algorithm, star seed)."""
    return _take(_load().lev_graph_barabasi_albert(int(n), int(m), int(seed) & (2**64 - 1)))


def from_edges(n: int, src, dst) -> CSRGraph:
    """CSR from an undirected edge list (each edge stored in both directions)."""
    src = np.ascontiguousarray(src, dtype=np.int32)
    dst = np.ascontiguousarray(dst, dtype=np.int32)
    if src.shape != dst.shape:
        raise ValueError("src and dst must have the same length")
    return _take(_load().lev_graph_from_edges(int(n), int(src.shape[0]), src.ctypes.data, dst.ctypes.data))


def save(path, g: CSRGraph, **arrays) -> None:
    """Write a graph (and optional per-node / per-edge arrays, e.g. delays) as an .npz of plain
    arrays: row_ptr int32[N+1], col_idx int32[E]."""
    np.savez(path, row_ptr=g.row_ptr, col_idx=g.col_idx, **arrays)


def load(path) -> tuple[CSRGraph, dict]:
    """Read a graph written by save(); returns (graph, the other arrays).  The CSR is checked:
    row_ptr starts at 0 and never decreases, and every column is a node."""
    with np.load(path, allow_pickle=False) as z:
        arrays = {k: z[k] for k in z.files}
    row_ptr = np.ascontiguousarray(arrays.pop("row_ptr"), dtype=np.int32)
    col_idx = np.ascontiguousarray(arrays.pop("col_idx"), dtype=np.int32)
    n = row_ptr.shape[0] - 1
    if (n < 0 or row_ptr[0] != 0 or (np.diff(row_ptr) < 0).any() or row_ptr[-1] != col_idx.shape[0]
            or (col_idx.size and (col_idx.min() < 0 or col_idx.max() >= n))):
        raise ValueError(f"{path} does not hold a valid CSR")
    return CSRGraph(row_ptr, col_idx), arrays


def permute(g: CSRGraph, perm) -> tuple[CSRGraph, np.ndarray]:
    """Relabel nodes (new node i is old node perm[i]); returns the new CSR, columns ascending
    within each row, and the edge map (old CSR index of each new entry) for per-edge arrays."""
    perm = np.ascontiguousarray(perm, dtype=np.int32)
    n = g.num_nodes
    if perm.shape != (n,):
        raise ValueError("perm must have one entry per node")
    row_ptr = np.empty(n + 1, dtype=np.int32)
    col_idx = np.empty(g.num_edges, dtype=np.int32)
    emap = np.empty(g.num_edges, dtype=np.int64)
    if _load().lev_graph_permute(n, g.row_ptr.ctypes.data, g.col_idx.ctypes.data, perm.ctypes.data,
                                 row_ptr.ctypes.data, col_idx.ctypes.data, emap.ctypes.data) != 0:
        raise ValueError(_load().lev_graph_last_error().decode())
    return CSRGraph(row_ptr, col_idx), emap


def degree_order(g: CSRGraph) -> np.ndarray:
    """Nodes by in-degree, highest first (ties by index): hubs get neighbouring labels, so their
    history lines share cache lines and the hub rows form the first row blocks."""
    return np.argsort(-g.degrees(), kind="stable").astype(np.int32)


def rcm_order(g: CSRGraph) -> np.ndarray:
    """Reverse Cuthill-McKee order of the (symmetric) graph, from scipy.sparse.csgraph: neighbours
    get nearby labels, which shortens gathers on graphs with bounded bandwidth."""
    from scipy.sparse import csr_matrix
    from scipy.sparse.csgraph import reverse_cuthill_mckee

    n = g.num_nodes
    m = csr_matrix((np.ones(g.num_edges, dtype=np.int8), g.col_idx, g.row_ptr), shape=(n, n))
    return np.ascontiguousarray(reverse_cuthill_mckee(m, symmetric_mode=True), dtype=np.int32)


def disjoint_union(graphs) -> tuple[CSRGraph, np.ndarray]:
    """Concatenate graphs into one block-diagonal CSR; returns (graph, node offsets[n_graphs+1])."""
    offs = np.zeros(len(graphs) + 1, dtype=np.int64)
    eoffs = np.zeros(len(graphs) + 1, dtype=np.int64)
    for i, g in enumerate(graphs):
        offs[i + 1] = offs[i] + g.num_nodes
        eoffs[i + 1] = eoffs[i] + g.num_edges
    if eoffs[-1] > np.iinfo(np.int32).max or offs[-1] > np.iinfo(np.int32).max:
        raise ValueError("union exceeds the int32 CSR range")
    row_ptr = np.empty(offs[-1] + 1, dtype=np.int32)
    col_idx = np.empty(eoffs[-1], dtype=np.int32)
    for i, g in enumerate(graphs):
        row_ptr[offs[i]:offs[i + 1]] = g.row_ptr[:-1] + eoffs[i]
        col_idx[eoffs[i]:eoffs[i + 1]] = g.col_idx + offs[i]
    row_ptr[-1] = eoffs[-1]
    return CSRGraph(row_ptr, col_idx), offs
