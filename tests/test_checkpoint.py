"""Checkpoint files (leviathan_v1_amd.checkpoint): plain-array .npz round trip and graph guard."""
import numpy as np
import pytest

from leviathan_v1_amd import checkpoint


def _state(N=50, E=300, D=7, nets=2, seed=0):
    rng = np.random.default_rng(seed)
    return {"theta": rng.random(N, dtype=np.float32), "theta_hat": rng.random(N, dtype=np.float32),
            "omega_hat": rng.random(N, dtype=np.float32), "weights": rng.random(E, dtype=np.float32),
            "history": rng.random((D, N), dtype=np.float32), "gains": rng.random(nets),
            "step": np.int64(123), "N": np.int64(N), "E": np.int64(E), "max_delay": np.int64(D)}


def test_round_trip_is_exact(tmp_path):
    s = _state()
    rp, ci, dl = np.arange(51, dtype=np.int32), np.zeros(300, np.int32), np.ones(300, np.uint8)
    p = tmp_path / "c.npz"
    checkpoint.save(p, s, rp, ci, dl)
    t = checkpoint.load(p, rp, ci, dl)
    for k in checkpoint.FIELDS:
        assert np.array_equal(np.asarray(s[k]), t[k]), k
        assert np.asarray(s[k]).dtype == t[k].dtype, k


def test_other_graph_is_refused(tmp_path):
    s = _state()
    rp, ci, dl = np.arange(51, dtype=np.int32), np.zeros(300, np.int32), np.ones(300, np.uint8)
    p = tmp_path / "c.npz"
    checkpoint.save(p, s, rp, ci, dl)
    dl2 = dl.copy()
    dl2[7] = 2
    with pytest.raises(ValueError):
        checkpoint.load(p, rp, ci, dl2)


def test_incomplete_state_is_refused(tmp_path):
    s = _state()
    del s["history"]
    with pytest.raises(ValueError):
        checkpoint.save(tmp_path / "c.npz", s)


def test_digest_ignores_integer_types(tmp_path):
    """The same graph held as int64 row pointers (scipy's indptr) and int32 columns digests
    the same, so its checkpoint loads; values outside the engine's types are refused."""
    s = _state()
    rp, ci, dl = np.arange(51, dtype=np.int32), np.zeros(300, np.int32), np.ones(300, np.uint8)
    assert checkpoint.graph_digest(rp, ci, dl) == checkpoint.graph_digest(rp.astype(np.int64), ci.astype(np.int64),
                                                                         dl.astype(np.int32))
    p = tmp_path / "c.npz"
    checkpoint.save(p, s, rp.astype(np.int64), ci, dl)
    checkpoint.load(p, rp, ci, dl)
    with pytest.raises(ValueError):
        checkpoint.graph_digest(rp.astype(np.int64) + 2**40, ci, dl)
