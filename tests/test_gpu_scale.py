"""The full-size step: Barabasi-Albert N = 10^7, m = 10 (2.0e8 incoming edges, in-degrees
up to ~10^4), delays uniform in [1, 50], FP32 and uint16 history.

  * per-step phases and weights of each form within 1e-5 relative of one FP64 oracle step
    taken from the engine's exported state (the bar of the other GPU tests), and the
    step's r within 1e-6 of the FP64 order parameter of the phases it wrote;
  * 2,000 steps of the uint16 form (the schedule the autotuner picks) against 2,000 steps of
    the plain FP32 drop-in from the same inputs: windowed (200-step) mean r and r variance
    within 5e-3.  An FP64 run of 4e11 edge-steps is out of reach of the CPU oracle, so
    the uint16 form's windowed r is checked against the FP64 oracle at N = 10^5
    (test_gpu_step) and against the FP32 form here.
"""
import numpy as np
import pytest

from gpu_helpers import oracle_from_engine, say, theta_rel_err, w_rel_err, windowed

pytestmark = pytest.mark.gpu

DT = 0.05
TOL = 1e-5


@pytest.fixture(scope="module")
def ba10m():
    from leviathan_v1_amd import workloads

    return workloads.barabasi_albert_workload(10_000_000, 10, 50, seed=0)


def make(wl, **kw):
    from leviathan_v1_amd import LeviathanEngine

    return LeviathanEngine(*wl.engine_args(), **kw)


@pytest.mark.parametrize("history,schedule", [("f32", "direct"), ("u16", "binned"), ("u16", "staged")])
def test_ba10m_one_step_vs_fp64_oracle(ba10m, history, schedule):
    say("ba10m one step", history, schedule)
    eng = make(ba10m, history=history, schedule=schedule)
    eng.step_n(DT, 20)
    o = oracle_from_engine(eng, ba10m)
    o.step(DT)
    r = eng.step(DT)
    th = eng.get_theta()
    et, ew = theta_rel_err(th, o.theta), w_rel_err(eng.get_weights(), o.w)
    # r of the step (10^5 row-block partials, summed in chunks, then the chunk sums) against
    # the FP64 order parameter of the phases the step wrote
    r_np = float(np.abs(np.exp(1j * th.astype(np.float64)).mean()))
    say("theta rel err", et, "w rel err", ew, "r", r, "r of phases", r_np, "blocks", eng.num_blocks)
    eng.close()
    assert et < TOL and ew < TOL
    assert abs(r - r_np) < 1e-6


def test_ba10m_u16_windowed_r_matches_fp32(ba10m):
    say("ba10m 2000 steps u16 vs f32")
    u = make(ba10m, history="u16", schedule="auto")
    u.autotune(DT, steps=2)
    r16 = u.step_n(DT, 2000)
    name = u.schedule_name
    u.close()
    f = make(ba10m)
    r32 = f.step_n(DT, 2000)
    f.close()
    assert np.isfinite(r16).all() and np.isfinite(r32).all()
    assert 0 <= r16.min() and r16.max() <= 1
    m16, v16 = windowed(r16, 200, 200)
    m32, v32 = windowed(r32, 200, 200)
    say(name, "max |mean r diff|", float(np.abs(m16 - m32).max()), "max |var diff|", float(np.abs(v16 - v32).max()))
    assert np.abs(m16 - m32).max() < 5e-3
    assert np.abs(v16 - v32).max() < 5e-3
