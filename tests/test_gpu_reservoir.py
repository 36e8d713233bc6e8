"""GPU parity of the reservoir path: feature capture, FP64 moment kernels, the
BF16 MFMA Gram and the refined ridge solve, against CPU FP64 references."""
import numpy as np
import pytest

import oracle
from gpu_helpers import say

pytestmark = pytest.mark.gpu

DT = 0.05


@pytest.fixture(scope="module")
def res():
    from leviathan_v1_amd import workloads
    from leviathan_v1_amd.reservoir import LeviathanReservoir, multisine_input

    wl = workloads.watts_strogatz_workload(5000, 20, 0.1, 50, seed=9)
    rng = np.random.default_rng(9)
    nodes = rng.choice(wl.N, 150, replace=False).astype(np.int32)  # F = 300: not a multiple of 128
    b = rng.uniform(-1, 1, wl.N).astype(np.float32)
    u = multisine_input(3200, seed=9)
    r = LeviathanReservoir(wl, nodes, b, max_rows=3200)
    r.run(u[:1600])
    r.run(u[1600:])
    return {"res": r, "wl": wl, "nodes": nodes, "b": b, "u": u}


def test_capture_rows_are_cos_sin_of_phases(res):
    """Row t holds cos/sin of the feature nodes' phases after step t; the capture does not
    change the dynamics (same r as plain driven steps)."""
    from leviathan_v1_amd import LeviathanEngine

    say("reservoir capture")
    r, wl, nodes, b, u = res["res"], res["wl"], res["nodes"], res["b"], res["u"]
    assert r.rows == 3200
    th = r.engine.get_theta()[nodes].astype(np.float64)
    last = r.features(3199, 1)[0]
    assert np.abs(last[:150] - np.cos(th)).max() < 2e-6
    assert np.abs(last[150:] - np.sin(th)).max() < 2e-6
    e = LeviathanEngine(*wl.engine_args())
    e.set_drive(b)
    e.step_n_driven(DT, u[:1600])
    e.step_n_driven(DT, u[1600:])
    assert np.array_equal(e.get_theta(), r.engine.get_theta())
    with pytest.raises(RuntimeError):
        r.run(u[:1])  # max_rows reached


def test_fp64_moment_kernels(res):
    say("colsum / matvec")
    r = res["res"]
    X = r.features(200, 2000).astype(np.float64)
    w = np.random.default_rng(1).standard_normal(2000)
    cs = r.engine.features_colsum(200, 2000)
    assert np.abs(cs - X.sum(axis=0)).max() < 1e-9 * 2000
    cw = r.engine.features_colsum(200, 2000, w)
    assert np.abs(cw - X.T @ w).max() < 1e-9 * 2000
    mean = X.mean(axis=0)
    v = np.random.default_rng(2).standard_normal(X.shape[1])
    mv = r.engine.features_matvec(200, 2000, mean, v)
    assert np.abs(mv - (X - mean) @ v).max() < 1e-9 * X.shape[1]


def test_mfma_gram_matches_fp64_gram_of_bf16_inputs(res):
    """The matrix-core Gram equals the FP64 Gram of the same BF16-rounded centred features
    up to FP32 accumulation error (checks the MFMA operand and accumulator layouts,
    the tile mirroring and the padding of F = 300 to 384 and T = 2,900 to 2,944)."""
    say("MFMA gram")
    r = res["res"]
    row0, n = 300, 2900
    X = r.features(row0, n).astype(np.float64)
    mean = X.mean(axis=0)
    G, ms = r.engine.features_gram(row0, n, mean)
    Xb = oracle.bf16_round((X - mean).astype(np.float32)).astype(np.float64)
    ref = Xb.T @ Xb
    scale = np.sqrt(np.outer(np.diag(ref), np.diag(ref)))
    assert np.abs(G - ref).max() / scale.max() < 1e-5
    assert np.array_equal(G, G.T)
    # and it approximates the exact centred Gram to BF16 accuracy
    exact = (X - mean).T @ (X - mean)
    assert np.abs(G - exact).max() / scale.max() < 2e-2


def test_split_gram_matches_fp64_gram_of_hi_lo_parts(res):
    """terms=3: the Gram of split BF16 inputs equals hi hi^T + hi lo^T + lo hi^T in FP64 up to
    FP32 accumulation, and approximates the exact centred Gram ~2^8 times closer than one BF16
    product (F = 300 padded to 384 columns, T = 2,900 rows padded to 2,944)."""
    say("split gram")
    r = res["res"]
    row0, n = 300, 2900
    X = r.features(row0, n).astype(np.float64)
    mean = X.mean(axis=0)
    G, ms = r.engine.features_gram(row0, n, mean, terms=3)
    xc = (X - mean).astype(np.float32)
    hi = oracle.bf16_round(xc)
    lo = oracle.bf16_round(xc - hi)
    H, L = hi.astype(np.float64), lo.astype(np.float64)
    ref = H.T @ H + H.T @ L + L.T @ H
    scale = np.sqrt(np.outer(np.diag(ref), np.diag(ref))).max()
    assert np.abs(G - ref).max() / scale < 1e-5
    assert np.array_equal(G, G.T)
    exact = (X - mean).T @ (X - mean)
    err3 = np.abs(G - exact).max() / scale
    G1, _ = r.engine.features_gram(row0, n, mean, terms=1)
    err1 = np.abs(G1 - exact).max() / scale
    say("split gram error", err3, "bf16 gram error", err1)
    assert err3 < 2e-5 and err3 < err1 / 20


@pytest.mark.parametrize("terms", [1, 0])
@pytest.mark.parametrize("alpha", [1e-4, 1.0])
def test_device_cholesky_solves_the_shifted_gram(res, alpha, terms):
    """features_factor / features_precond: z solves (G + (alpha + shift) I) z = r to FP64 accuracy,
    with G the matrix-core Gram features_gram returns for the same rows (terms 1) or the FP64 Gram
    of the centred features (terms 0), F = 300; the shift is the smallest in 0, 1e-7 s, 1e-6 s, ...
    that factorises (numpy's Cholesky agrees)."""
    say("device cholesky", alpha, terms)
    r = res["res"]
    row0, n = 300, 2900
    X = r.features(row0, n).astype(np.float64)
    mean = X.mean(axis=0)
    if terms == 1:
        G, _ = r.engine.features_gram(row0, n, mean)
        G = G.astype(np.float64)
    else:
        G = (X - mean).T @ (X - mean)
    shift, ms = r.engine.features_factor(row0, n, mean, alpha, terms=terms)
    A = G + (alpha + shift) * np.eye(G.shape[0])
    np.linalg.cholesky(A)  # factorisable at the shift found
    if shift > 0:
        with pytest.raises(np.linalg.LinAlgError):
            np.linalg.cholesky(G + (alpha + shift / 10) * np.eye(G.shape[0]))
    rhs = np.random.default_rng(9).standard_normal(G.shape[0])
    z = r.engine.features_precond(rhs)
    ref = np.linalg.solve(A, rhs)
    cond = np.linalg.cond(A)
    assert np.linalg.norm(A @ z - rhs) / np.linalg.norm(rhs) < 1e-13 * cond
    assert np.abs(z - ref).max() / np.abs(ref).max() < 1e-14 * cond


def test_ridge_fit_matches_fp64_oracle(res):
    """Preconditioned FP64 refinement reaches the FP64 ridge solution of the same features."""
    say("ridge fit")
    r, u = res["res"], res["u"]
    lag, row0 = 10, 200
    n = 3200 - row0 - lag
    y = u[row0 + lag: row0 + lag + n].astype(np.float64)
    for alpha, terms in ((1e-4, None), (1.0, 1), (1e-4, 3), (1e-4, 0)):
        fit = r.fit(y, row0, alpha=alpha, tol=1e-12, gram_terms=terms)
        X = r.features(row0, n)
        w_ref, c_ref = oracle.ridge_fp64(X, y, alpha)
        pred = r.predict(fit, row0, n)
        pred_ref = X.astype(np.float64) @ w_ref + c_ref
        say("alpha", alpha, "iterations", fit.iterations, "rel residual", fit.rel_residual, "shift", fit.shift,
            "prediction diff", float(np.abs(pred - pred_ref).max()))
        assert fit.rel_residual < 1e-8
        assert np.abs(pred - pred_ref).max() < 1e-6 * (np.abs(pred_ref).max() + 1)
