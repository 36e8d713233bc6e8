"""The allocation-free moments entry point (leviathan_features_colsum_noalloc), called
through the C-ABI, against leviathan_features_colsum on the same captured features: equal
bit for bit (the same kernels), on two calls in a row, the second with a smaller input, with
and without weights; sizes beyond its stated maximum are refused with an error."""
import ctypes

import numpy as np
import pytest

from gpu_helpers import say

pytestmark = pytest.mark.gpu

DT = 0.05


def _colsum(lib, h, name, row0, n, w, F):
    out = np.full(F, np.nan)
    wp = None if w is None else ctypes.c_void_p(w.ctypes.data)
    rc = getattr(lib, name)(h, row0, n, wp, ctypes.c_void_p(out.ctypes.data))
    return rc, out


def test_colsum_noalloc_matches_colsum():
    from leviathan_v1_amd import LeviathanEngine, _lib, workloads

    lib = _lib.lib()
    wl = workloads.watts_strogatz_workload(3000, 20, 0.1, 50, seed=21)
    eng = LeviathanEngine(*wl.engine_args())
    rng = np.random.default_rng(21)
    nodes = np.sort(rng.choice(wl.N, 700, replace=False)).astype(np.int32)  # F = 1,400
    eng.features_setup(nodes, 2400)
    eng.step_n_features(DT, 2400)
    h = eng._h
    F = 1400
    for row0, n, weighted in ((100, 2300, True), (50, 400, True), (0, 2400, False), (7, 9, True)):
        w = rng.standard_normal(n) if weighted else None
        say("colsum noalloc", row0, n, weighted)
        rc1, a = _colsum(lib, h, "leviathan_features_colsum", row0, n, w, F)
        rc2, b = _colsum(lib, h, "leviathan_features_colsum_noalloc", row0, n, w, F)
        assert rc1 == 0 and rc2 == 0, _lib.last_error()
        assert np.array_equal(a, b)
        X = eng.features_get(row0, n).astype(np.float64)
        ref = (X * (w[:, None] if weighted else 1.0)).sum(axis=0)
        assert np.abs(b - ref).max() < 1e-9 * max(1.0, np.abs(ref).max())
    eng.close()


def test_colsum_noalloc_limits():
    from leviathan_v1_amd import LeviathanEngine, _lib, workloads

    lib = _lib.lib()
    wl = workloads.watts_strogatz_workload(5000, 20, 0.1, 50, seed=22)
    # more features than the scratch holds (2 * 4,100 > 8,192): refused, the allocating call works
    big = LeviathanEngine(*wl.engine_args())
    big.features_setup(np.arange(4100, dtype=np.int32), 10)
    big.step_n_features(DT, 10)
    rc, _ = _colsum(lib, big._h, "leviathan_features_colsum_noalloc", 0, 10, None, 8200)
    assert rc == -1 and "allocation-free maximum" in _lib.last_error()
    rc, _ = _colsum(lib, big._h, "leviathan_features_colsum", 0, 10, None, 8200)
    assert rc == 0
    big.close()
    # more weighted rows than the scratch holds: refused; unweighted they are fine
    small = LeviathanEngine(*workloads.watts_strogatz_workload(200, 4, 0.1, 5, seed=23).engine_args())
    small.features_setup(np.arange(8, dtype=np.int32), 65_600)
    small.step_n_features(DT, 65_600)
    w = np.ones(65_600)
    rc, _ = _colsum(lib, small._h, "leviathan_features_colsum_noalloc", 0, 65_600, w, 16)
    assert rc == -1 and "allocation-free maximum" in _lib.last_error()
    rc, a = _colsum(lib, small._h, "leviathan_features_colsum_noalloc", 0, 65_600, None, 16)
    assert rc == 0
    rc, b = _colsum(lib, small._h, "leviathan_features_colsum", 0, 65_600, None, 16)
    assert rc == 0 and np.array_equal(a, b)
    small.close()
