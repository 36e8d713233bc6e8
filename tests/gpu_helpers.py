"""Shared helpers of the GPU parity tests: move an engine's state into the oracle."""
import numpy as np

from oracle import COracle, Params, circ_err

TWO_PI = 2 * np.pi


def say(*msg):
    print("[case]", *msg, flush=True)


def oracle_from_engine(eng, wl, params=None, dtype=np.float64, drive_b=None, parallel=False):
    """An oracle holding exactly the engine's current state (phases, predictor,
    weights, every history slot, gain), so one oracle step from it isolates the
    error of one GPU step.  parallel: the OpenMP FP64 restatement (same phases and
    weights as the sequential one; for the 10^7-node graph)."""
    theta = eng.get_theta()
    thh, omh = eng.get_predictor()
    o = COracle(wl.N, wl.max_delay, wl.graph.row_ptr, wl.graph.col_idx, wl.delays, eng.get_weights(), theta, thh,
                wl.omega, params=params or Params(), dtype=dtype, parallel=parallel)
    o.omega_hat[:] = omh
    B = o.B
    for d in range(wl.max_delay + 1):
        o.buf[:, (-d) % B] = eng.get_history(d)
    o.buf_idx = 0
    o.g = o.dtype.type(eng.get_gain())
    if drive_b is not None:
        o.drive_b = np.asarray(drive_b, dtype=dtype)
    return o


def theta_rel_err(a, b):
    """Phase error on the circle, relative to the 2*pi range of theta."""
    return float(circ_err(a, b).max() / TWO_PI)


def w_rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float((np.abs(a - b) / np.maximum(np.abs(b), 1e-3)).max())


def windowed(r, warm, win):
    r = np.asarray(r, np.float64)[warm:]
    k = len(r) // win
    r = r[: k * win].reshape(k, win)
    return r.mean(axis=1), r.var(axis=1)
