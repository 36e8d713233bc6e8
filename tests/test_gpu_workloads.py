"""The measured configurations at their own sizes, and the Observatory against the oracle.

  * Batched ensemble (the counterpart of experiments/phase_sweep.py:25-59): 256 Watts-Strogatz
    networks (k=20, p=0.1), sizes log-uniform in [2,000, 200,000], gains K on linspace(0, 4),
    controller off, uint16 history, the autotuned schedule, 5,000 steps in 100-step graph
    chunks.  Every network's r is finite and in [0, 1], the last r of every network equals the
    FP64 order parameter of its final phases, and the three smallest networks' windowed r
    statistics match their FP64 oracle runs.
  * Reservoir readout (the counterpart of python/reservoir.py:21-94): N=50,000, 4,096 feature
    nodes (8,192 features), 50,000 driven steps, ridge 1e-4 predicting u(t+10).  A sampled
    block of the matrix-core Gram against the FP64 Gram of the same BF16 inputs, and the
    refined ridge solution against the FP64 Gram/solve of the same features.
  * LeviathanObservatory (python/leviathan_h100.py:82-93, :157-196): step, step_n and
    run_with_stimulus against the FP64 oracle given the same phase kick.
"""
import numpy as np
import pytest

import oracle
from gpu_helpers import say, theta_rel_err, windowed

pytestmark = pytest.mark.gpu

DT = 0.05


def test_ensemble_256_networks_5000_steps():
    from leviathan_v1_amd.ensemble import LeviathanEnsemble, build_workloads, sweep_spec

    spec = sweep_spec(256)
    wls = build_workloads(spec)
    say("ensemble 256 networks", "N", sum(w.N for w in wls), "E", sum(w.E for w in wls))
    ens = LeviathanEnsemble(wls, spec.gains, history="u16", schedule="auto", graph_chunk=100)
    res = ens.engine.autotune_all(DT, steps=10)
    say("autotune", ens.engine.schedule_name, "scatter steps", ens.engine.scatter_steps, res.get("plans"))
    r = ens.step_n(5000)
    assert r.shape == (5000, 256)
    assert np.isfinite(r).all() and r.min() >= 0.0 and r.max() <= 1.0
    th = ens.engine.get_theta().astype(np.float64)
    for k in range(256):
        t = th[ens.offsets[k]:ens.offsets[k + 1]]
        assert abs(float(np.abs(np.exp(1j * t).mean())) - float(r[-1, k])) < 1e-6, k
    tail = r[-1000:].mean(axis=0)
    say("r(K) over the last 1,000 steps, every 32nd network:", np.round(tail[::32], 3).tolist())
    ens.close()
    # The three smallest networks (N ~ 2,000) against their FP64 oracle runs (same gains,
    # controller off).  The dynamics are chaotic and at this size r fluctuates by ~1/sqrt(N)
    # ~ 0.02 on slow time scales, so two FP64 trajectories whose initial phases differ by the
    # uint16 history's quantisation (up to pi/65536 per node) also separate in windowed r.  The
    # bar is 5e-3 or twice that FP64 chaos floor (measured for each network), whichever is
    # larger.
    for k in np.argsort(spec.sizes)[:3]:
        prm = oracle.Params(g0=float(spec.gains[k]), gain_controller=False)
        args = wls[k].engine_args()
        ref = oracle.COracle(*args, params=prm, dtype=np.float64, parallel=True).run(DT, 5000)
        rng = np.random.default_rng(int(k))
        th_p = np.asarray(args[6], np.float64) + rng.uniform(-np.pi / 65536, np.pi / 65536, wls[k].N)
        alt = oracle.COracle(*args[:6], th_p, th_p, args[8], params=prm, dtype=np.float64, parallel=True).run(DT, 5000)
        m, v = windowed(r[:, k], 1000, 1000)
        mr, vr = windowed(ref, 1000, 1000)
        ma, va = windowed(alt, 1000, 1000)
        floor_m, floor_v = float(np.abs(ma - mr).max()), float(np.abs(va - vr).max())
        dm, dv = float(np.abs(m - mr).max()), float(np.abs(v - vr).max())
        say("net", int(k), "N", wls[k].N, "K", float(spec.gains[k]), "max |mean r diff|", dm, "max |var diff|", dv,
            "FP64 chaos floor", floor_m, floor_v)
        assert dm < max(5e-3, 2 * floor_m), k
        assert dv < max(5e-3, 2 * floor_v), k


@pytest.fixture(scope="module")
def reservoir50k():
    from leviathan_v1_amd import workloads
    from leviathan_v1_amd.reservoir import LeviathanReservoir, multisine_input

    N, steps, n_nodes, seed = 50_000, 50_000, 4096, 0
    wl = workloads.watts_strogatz_workload(N, 20, 0.1, 50, seed=seed)
    rng = np.random.Generator(np.random.PCG64(seed + 7))
    nodes = np.sort(rng.choice(N, n_nodes, replace=False)).astype(np.int32)
    b = rng.uniform(-1.0, 1.0, N).astype(np.float32)
    u = multisine_input(steps, seed=seed + 11)
    res = LeviathanReservoir(wl, nodes, b, max_rows=steps)
    rr = res.run(u)
    assert np.isfinite(rr).all()
    yield {"res": res, "u": u}
    res.close()


def test_reservoir_cfg_gram_block_vs_fp64(reservoir50k):
    """F = 8,192 features over 39,192 training rows: sampled 256 x 256 blocks of the BF16
    matrix-core Gram (on the diagonal, off it, and at the far corner) equal the FP64 Gram of
    the same BF16-rounded centred features up to FP32 accumulation error."""
    res = reservoir50k["res"]
    row0, n = 1000, 39_192
    X = res.features(row0, n)
    mean = res.engine.features_colsum(row0, n) / n
    G, ms = res.engine.features_gram(row0, n, mean)
    say("gram F", G.shape[0], "kernel ms", ms)
    Xb = oracle.bf16_round((X.astype(np.float64) - mean).astype(np.float32)).astype(np.float64)
    dg = np.einsum("ij,ij->j", Xb, Xb)
    for i0, j0 in ((0, 0), (0, 4096), (4096, 4352), (7936, 7936), (256, 7936)):
        ref = Xb[:, i0:i0 + 256].T @ Xb[:, j0:j0 + 256]
        scale = np.sqrt(np.outer(dg[i0:i0 + 256], dg[j0:j0 + 256]))
        err = float((np.abs(G[i0:i0 + 256, j0:j0 + 256] - ref) / scale).max())
        say("block", i0, j0, "rel err", err)
        assert err < 1e-5, (i0, j0)
    assert np.array_equal(G, G.T)


def test_reservoir_cfg_ridge_vs_fp64(reservoir50k):
    """The readout (BF16 Gram preconditioner, FP64 refinement on the device) reaches the FP64
    ridge solution of the same features: relative solution error below 1e-10."""
    res, u = reservoir50k["res"], reservoir50k["u"]
    lag, washout, alpha = 10, 1000, 1e-4
    rows = 50_000 - washout - lag
    n_train = int(0.8 * rows)
    y = u[washout + lag: washout + lag + rows].astype(np.float64)
    fit = res.fit(y[:n_train], washout, alpha=alpha)
    X = res.features(washout, n_train)
    w64, c64 = oracle.ridge_fp64(X, y[:n_train], alpha)
    err = float(np.linalg.norm(fit.w - w64) / np.linalg.norm(w64))
    say("ridge iterations", fit.iterations, "rel residual", fit.rel_residual, "solution rel err", err,
        "intercept", fit.intercept, c64)
    assert err < 1e-10
    assert abs(fit.intercept - c64) < 1e-8 * (abs(c64) + 1)


def _kick(theta, nodes, strength):
    th = np.array(theta, dtype=np.float64)
    th[nodes] += strength
    return th % (2 * np.pi)


def test_observatory_step_and_stimulus_vs_oracle():
    """LeviathanObservatory over a seeded workload: step() and step_n() against the FP64 oracle
    per step (1e-5 relative, 10 steps), then run_with_stimulus with the phase kick (which
    erases the delay history, leviathan_h100.py:183-190) applied to the oracle too: per-step
    phases and r for the first steps after the kick, and windowed r over 3,000 steps."""
    from leviathan_v1_amd import LeviathanObservatory, workloads

    wl = workloads.watts_strogatz_workload(2000, 20, 0.2, 50, seed=42, reference_defaults=True)
    obs = LeviathanObservatory(workload=wl)
    o = oracle.COracle(*wl.engine_args(), dtype=np.float64)
    for t in range(5):
        r = obs.step()
        rr = o.step(DT)
        assert theta_rel_err(obs.get_phase_snapshot(), o.theta) < 1e-5, t
        assert abs(r - rr) < 1e-5, t
    r = obs.step_n(5)
    rr = o.run(DT, 5)
    assert theta_rel_err(obs.get_phase_snapshot(), o.theta) < 1e-5
    assert np.abs(r - rr).max() < 1e-5
    assert obs.statistics()["steps"] == 10 and np.array_equal(obs.r_history[-5:], r)

    # a short stimulus run: kick 10 nodes by 2 rad after step 3 of 8
    nodes = list(range(10))
    obs2 = LeviathanObservatory(workload=wl)
    o2 = oracle.COracle(*wl.engine_args(), dtype=np.float64)
    r_obs = obs2.run_with_stimulus(num_steps=8, stimulus_start=3, stimulus_nodes=nodes, stimulus_strength=2.0)
    r_ref = []
    for i in range(8):
        r_ref.append(o2.step(DT))
        if i == 3:
            o2.set_theta(_kick(o2.theta, nodes, 2.0))
    assert theta_rel_err(obs2.get_phase_snapshot(), o2.theta) < 1e-5
    assert np.abs(np.asarray(r_obs, np.float64) - r_ref).max() < 1e-5
    # the kick erased the delay history: delay D still reads the kicked phases, in both
    assert theta_rel_err(obs2.engine.get_history(wl.max_delay), o2.history(wl.max_delay)) < 1e-5

    # a long one: 3,000 steps with the kick at step 100, windowed r
    obs3 = LeviathanObservatory(workload=wl)
    o3 = oracle.COracle(*wl.engine_args(), dtype=np.float64)
    r_obs = obs3.run_with_stimulus(num_steps=3000, stimulus_start=100, stimulus_nodes=nodes)
    r_ref = np.empty(3000)
    r_ref[:101] = o3.run(DT, 101)
    o3.set_theta(_kick(o3.theta, nodes, 2.0))
    r_ref[101:] = o3.run(DT, 2899)
    m, v = windowed(r_obs, 500, 500)
    mr, vr = windowed(r_ref, 500, 500)
    say("observatory stimulus max |mean r diff|", float(np.abs(m - mr).max()), "max |var diff|",
        float(np.abs(v - vr).max()))
    assert np.abs(m - mr).max() < 5e-3 and np.abs(v - vr).max() < 5e-3
    for x in (obs, obs2, obs3):
        x.close()
