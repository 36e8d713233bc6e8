"""Pins the CPU oracle (oracle/) against closed forms and invariants.

The reference holds no golden vectors or known-answer tests for the step
(it has no tests at all), so every part of the oracle is pinned here by a
case whose answer the mathematics fixes, chosen so that a plausible slip (a
dropped term, a wrong sign, an off-by-one delay slot) fails at least one.
Both FP64 restatements (NumPy and C) are checked by every test.
"""
import numpy as np
import pytest

import oracle
from oracle import COracle, NumpyOracle, Params, circ_err

TWO_PI = 2 * np.pi
DT = 0.05


def tiny(kind, N, edges, theta, omega=None, theta_hat=None, max_delay=5, **pk):
    """edges: list of (src, dst, delay, weight); CSR of incoming edges."""
    rows = [[] for _ in range(N)]
    for s, d, dl, w in edges:
        rows[d].append((s, dl, w))
    row_ptr = np.zeros(N + 1, np.int32)
    col, dly, wts = [], [], []
    for i in range(N):
        rows[i].sort()
        for s, dl, w in rows[i]:
            col.append(s)
            dly.append(dl)
            wts.append(w)
        row_ptr[i + 1] = len(col)
    theta = np.asarray(theta, np.float64)
    omega = np.zeros(N) if omega is None else np.asarray(omega, np.float64)
    theta_hat = theta.copy() if theta_hat is None else np.asarray(theta_hat, np.float64)
    params = Params(**pk)
    args = (N, max_delay, row_ptr, np.array(col, np.int32), np.array(dly, np.uint8), np.array(wts, np.float64),
            theta, theta_hat, omega)
    if kind == "numpy":
        return NumpyOracle(*args, params=params)
    return COracle(*args, params=params, dtype=np.float64, parallel=kind == "c64p")


KINDS = ["numpy", "c64", "c64p"]


@pytest.fixture(scope="module", autouse=True)
def _built(oracle_lib):
    return oracle_lib


@pytest.mark.parametrize("kind", KINDS)
def test_free_rotation_closed_form(kind):
    """g=0, alpha=0: theta(t) = (theta0 + omega*t*dt) mod 2pi (integrate + wrap)."""
    th0 = np.array([0.1, 3.0, 6.2, 1.0])
    om = np.array([0.7, -1.3, 2.0, 0.0])
    o = tiny(kind, 4, [], th0, om, g0=0.0, gain_controller=False, alpha_eps=0.0)
    for t in range(1, 301):
        o.step(DT)
        want = np.remainder(th0 + om * t * DT, TWO_PI)
        assert circ_err(o.theta, want).max() < 1e-11, t
        assert np.all(o.theta >= 0) and np.all(o.theta < TWO_PI)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("d", [1, 3, 5])
def test_delay_ring_semantics(kind, d):
    """Edge 0->1 with delay d: row 1 sees theta_0(t-d), or theta_0(0) while t < d."""
    D = 5
    th0 = np.array([0.4, 2.0])
    om = np.array([1.1, 0.0])
    w = 0.3
    o = tiny(kind, 2, [(0, 1, d, w)], th0, om, max_delay=D, g0=0.0, gain_controller=False, alpha_eps=0.0,
             eta=0.0, lam=0.0)
    for t in range(0, 40):
        th1 = o.theta[1]
        o.step(DT)
        src = th0[0] + om[0] * max(t - d, 0) * DT
        want = w * np.sin(src - th1)
        assert abs(o.S[1] - want) < 1e-12, (t, o.S[1], want)
        # the ring holds theta_0(t+1-k) for k = 0..D
        for k in range(D + 1):
            assert circ_err(o.history(k)[0], th0[0] + om[0] * max(t + 1 - k, 0) * DT) < 1e-11


@pytest.mark.parametrize("kind", KINDS)
def test_coupling_known_values(kind):
    """S_i = sum_j w_ij sin(theta_j(t-d) - theta_i): sources at pi/2 and pi/6 of a node at 0."""
    o = tiny(kind, 3, [(1, 0, 1, 0.3), (2, 0, 2, 0.5)], [0.0, np.pi / 2, np.pi / 6], g0=0.0,
             gain_controller=False)
    o.step(DT)
    assert abs(o.S[0] - (0.3 * 1.0 + 0.5 * 0.5)) < 1e-12
    # uses the PRE-update weight (leviathan_kernel.cu:154) -- weights changed after the step
    assert o.w[0] != 0.3


@pytest.mark.parametrize("kind", KINDS)
def test_coupling_enters_phase_with_gain(kind):
    """theta_i(1) = theta_i + dt*(omega_i + g*S_i - alpha*eps_i) with eps = 0."""
    g = 1.7
    o = tiny(kind, 2, [(1, 0, 1, 0.8)], [1.0, 1.0 + np.pi / 2], [0.2, 0.0], g0=g, gain_controller=False)
    o.step(DT)
    assert abs(o.theta[0] - (1.0 + DT * (0.2 + g * 0.8))) < 1e-12


@pytest.mark.parametrize("kind", KINDS)
def test_prep_uses_old_omega_hat_and_eps_before_update(kind):
    """prep (:101-109): eps = sin(wrap(theta - theta_hat)); theta_hat += dt*omega_hat(old);
    omega_hat += kappa_w*dt*eps."""
    th = np.array([1.0, 5.9])
    thh = np.array([1.3, 0.2])  # second: difference 5.7 wraps to 5.7 - 2pi
    om = np.array([0.5, -0.25])
    o = tiny(kind, 2, [], th, om, theta_hat=thh, g0=0.0, gain_controller=False, kappa_w=0.01)
    o.step(DT)
    eps = np.sin(np.array([-0.3, 5.7 - TWO_PI]))
    assert np.allclose(o.eps, eps, atol=1e-14, rtol=0)
    assert np.allclose(o.theta_hat, thh + DT * om, atol=1e-14, rtol=0)
    assert np.allclose(o.omega_hat, om + 0.01 * DT * eps, atol=1e-14, rtol=0)
    # and eps feeds back into the phase with -alpha
    assert np.allclose(o.theta, np.remainder(th + DT * (om - 0.1 * eps), TWO_PI), atol=1e-13, rtol=0)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("eps_gap", [0.0, 0.9])
def test_hebbian_closed_form(kind, eps_gap):
    """Constant coupling angle Delta and constant eps: w_t = w* + (w0 - w*)(1 - lam*dt)^t,
    w* = eta*cos(Delta)*(1 - gamma*|eps|)/lam (clamped to [w_min, w_max])."""
    delta = np.arccos(0.1)
    w0 = 0.25
    th = np.array([0.0, delta])
    thh = th - np.array([eps_gap, 0.0])  # node 0's predictor lags by eps_gap
    o = tiny(kind, 2, [(1, 0, 2, w0)], th, theta_hat=thh, g0=0.0, gain_controller=False, alpha_eps=0.0,
             kappa_w=0.0)
    # theta_hat keeps moving with omega_hat = omega = 0, so eps stays sin(eps_gap)
    p = o.params
    gate = 1.0 - p.gamma * abs(np.sin(eps_gap))
    wstar = p.eta * 0.1 * gate / p.lam
    for t in range(1, 200):
        o.step(DT)
        want = wstar + (w0 - wstar) * (1 - p.lam * DT) ** t
        assert abs(o.w[0] - want) < 1e-12, (t, o.w[0], want)


@pytest.mark.parametrize("kind", KINDS)
def test_weight_clamp_saturates_at_w_max(kind):
    """Synchronised pair: w* = eta/lam = 10 > w_max, so w reaches w_max = 2 and stays."""
    o = tiny(kind, 2, [(1, 0, 1, 1.9)], [0.5, 0.5], g0=0.0, gain_controller=False)
    for _ in range(3000):
        o.step(DT)
    assert o.w[0] == 2.0


@pytest.mark.parametrize("kind", KINDS)
def test_nan_weight_guard(kind):
    """A NaN weight: after one step w = w_min and the row's phase is reset to 0 (:160-161, :189-190)."""
    o = tiny(kind, 3, [(1, 0, 1, 0.5), (2, 0, 1, 0.5), (0, 1, 1, 0.5)], [0.3, 1.0, 2.0], [0.1, 0.2, 0.3])
    w = o.w.copy()
    w[0] = np.nan
    o.reset_weights(w)
    o.step(DT)
    assert o.w[0] == o.params.w_min
    assert o.theta[0] == 0.0
    assert np.isfinite(o.w[1:]).all() and o.theta[1] != 0.0 and np.isfinite(o.theta).all()


@pytest.mark.parametrize("kind", KINDS)
def test_nan_phase_cascade(kind):
    """NaN phase at node 2 (set_theta also fills its history with NaN):
    its omega_hat turns NaN (prep has no guard), its theta is reset to 0, and every
    row reading it gets theta 0 and w_min on that edge."""
    o = tiny(kind, 4, [(2, 0, 1, 0.5), (1, 0, 1, 0.5), (3, 1, 1, 0.5), (0, 3, 1, 0.5)],
             [0.3, 1.0, 2.0, 4.0], [0.1, 0.2, 0.3, 0.4])
    th = o.get_theta()
    th[2] = np.nan
    o.set_theta(th)
    o.step(DT)
    assert np.isnan(o.omega_hat[2]) and np.isfinite(o.theta_hat[2])
    assert o.theta[2] == 0.0
    assert o.theta[0] == 0.0
    e_from_2 = 1  # row 0's columns sorted: [1, 2] -> the edge from node 2 is index 1
    assert o.w[e_from_2] == 0.0
    assert np.isfinite(o.theta[[1, 3]]).all() and o.theta[1] != 0.0
    o.step(DT)
    assert np.isnan(o.theta_hat[2])


@pytest.mark.parametrize("kind", KINDS)
def test_order_parameter_extremes(kind):
    """All phases equal: r = 1.  Evenly spaced 2*pi*k/N: r = 0."""
    N = 64
    o = tiny(kind, N, [], np.full(N, 1.234), g0=0.0, gain_controller=False)
    assert abs(o.step(DT) - 1.0) < 1e-12
    o = tiny(kind, N, [], TWO_PI * np.arange(N) / N, g0=0.0, gain_controller=False)
    assert o.step(DT) < 1e-12


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("phases, rconst", [("sync", 1.0), ("splay", 0.0)])
def test_gain_controller_closed_form(kind, phases, rconst):
    """Constant r: g_t = g* + (g0 - g*)(1 - mu*dt)^t, g* = g_rest + beta*(r0 - r)/mu."""
    N = 32
    th = np.full(N, 0.7) if phases == "sync" else TWO_PI * np.arange(N) / N
    o = tiny(kind, N, [], th, g0=1.5)
    p = o.params
    gstar = p.g_rest + p.beta * (p.r0 - rconst) / p.mu
    for t in range(1, 400):
        o.step(DT)
        want = gstar + (p.g0 - gstar) * (1 - p.mu * DT) ** t
        assert abs(o.g - want) < 1e-12


@pytest.mark.parametrize("kind", KINDS)
def test_gain_clamp(kind):
    """A strong pull below g_min clamps at g_min; above g_max clamps at g_max."""
    N = 16
    o = tiny(kind, N, [], np.full(N, 0.1), g0=0.2, beta=10.0)  # r = 1 > r0: pushes g down
    for _ in range(50):
        o.step(DT)
    assert o.g == o.params.g_min
    o = tiny(kind, N, [], TWO_PI * np.arange(N) / N, g0=4.9, beta=10.0)  # r = 0: pushes g up
    for _ in range(50):
        o.step(DT)
    assert o.g == o.params.g_max


@pytest.mark.parametrize("kind", KINDS)
def test_set_theta_resets_history(kind):
    o = tiny(kind, 3, [(1, 0, 2, 0.5), (0, 1, 3, 0.5)], [0.1, 0.2, 0.3], [1.0, 2.0, 3.0])
    for _ in range(7):
        o.step(DT)
    new = np.array([1.0, 2.0, 3.0])
    o.set_theta(new)
    for k in range(o.max_delay + 1):
        assert np.array_equal(o.history(k), new)


@pytest.mark.parametrize("kind", KINDS)
def test_input_drive_closed_form(kind):
    """The input drive omega_i(t) = omega_i + b_i*u_t (the reservoir's extension; the reference
    has no input term) with g = 0 and alpha = 0: theta_i(t) = wrap(theta_i(0) +
    dt*sum_{s<t}(omega_i + b_i*u_s)).  A drive read one step late, with the wrong sign, or
    applied to the predictor instead of the phase fails it (u varies from step to step and
    b has both signs)."""
    th0 = np.array([0.3, 2.5, 5.0, 1.0])
    om = np.array([0.4, -0.9, 1.5, 0.0])
    b = np.array([1.0, -0.5, 0.25, 2.0])
    u = np.sin(0.37 * np.arange(200)) + 0.1 * np.cos(1.3 * np.arange(200))
    o = tiny(kind, 4, [(0, 1, 2, 0.7), (2, 3, 1, 0.4)], th0, om, g0=0.0, gain_controller=False, alpha_eps=0.0,
             kappa_w=0.0)
    o.drive_b = b.copy()
    acc = th0.copy()
    for t in range(200):
        o.step(DT, u[t])
        acc = acc + DT * (om + b * u[t])
        assert circ_err(o.theta, np.remainder(acc, TWO_PI)).max() < 1e-11, t
    # with kappa_w = 0 omega_hat stays the undriven omega: the predictor does not see the drive
    assert np.allclose(o.theta_hat, th0 + 200 * DT * om, atol=1e-11, rtol=0)


def test_numpy_and_c_agree_with_drive(oracle_lib):
    """NumPy and both C restatements agree over 300 driven steps with plasticity, the gain
    controller and the coupling on."""
    from leviathan_v1_amd import workloads

    wl = workloads.small_plain(5)
    a = wl.engine_args()
    rng = np.random.default_rng(2)
    b = rng.uniform(-1, 1, wl.N)
    u = rng.normal(0, 1, 300)
    c = COracle(*a, dtype=np.float64)
    cp = COracle(*a, dtype=np.float64, parallel=True)
    n = NumpyOracle(*a)
    for o in (c, cp, n):
        o.drive_b = b.copy()
    rc = c.run(DT, 300, u)
    rp = cp.run(DT, 300, u)
    rn = n.run(DT, 300, u)
    assert np.abs(rc - rn).max() < 1e-12
    assert circ_err(c.theta, n.theta).max() < 1e-10
    assert np.abs(c.w - n.w).max() < 1e-12
    assert abs(c.g - n.g) < 1e-13
    # the OpenMP restatement: phases, weights and gain bit-identical, r summed in blocks
    assert np.array_equal(c.theta, cp.theta) and np.array_equal(c.w, cp.w) and c.g == cp.g
    assert np.abs(rc - rp).max() < 1e-14
    # and the drive matters: the same run undriven ends elsewhere
    c0 = COracle(*a, dtype=np.float64)
    c0.run(DT, 300)
    assert circ_err(c0.theta, c.theta).max() > 1e-2


def test_parallel_fp64_oracle_bit_identical_on_ba(oracle_lib):
    """The OpenMP FP64 restatement (used for the 10^7-node fixture) equals the sequential
    one bit for bit in phases, weights and gain on a heavy-tailed graph (rows of hundreds of
    edges, so the dynamic row schedule splits them across threads)."""
    from leviathan_v1_amd import workloads

    wl = workloads.barabasi_albert_workload(20_000, 10, 50, seed=4)
    a = wl.engine_args()
    c = COracle(*a, dtype=np.float64)
    cp = COracle(*a, dtype=np.float64, parallel=True)
    rc = c.run(DT, 40)
    rp = cp.run(DT, 40)
    assert np.array_equal(c.theta, cp.theta) and np.array_equal(c.w, cp.w) and c.g == cp.g
    assert np.abs(rc - rp).max() < 1e-14


def test_numpy_and_c_fp64_agree_on_small_workload(oracle_lib):
    """Two independent FP64 restatements of the whole step agree over 300 steps."""
    from leviathan_v1_amd import workloads

    wl = workloads.small_plain(3)
    a = wl.engine_args()
    c = COracle(*a, dtype=np.float64)
    n = NumpyOracle(*a)
    rc = c.run(DT, 300)
    rn = n.run(DT, 300)
    assert np.abs(rc - rn).max() < 1e-12
    assert circ_err(c.theta, n.theta).max() < 1e-10
    assert np.abs(c.w - n.w).max() < 1e-12
    assert abs(c.g - n.g) < 1e-13


def test_fp32_port_tracks_fp64_short_horizon(oracle_lib):
    """The FP32 CPU baseline stays within 1e-5 (relative to 2*pi) of FP64 for 10 steps."""
    from leviathan_v1_amd import workloads

    wl = workloads.small_plain(0)
    a = wl.engine_args()
    c64 = COracle(*a, dtype=np.float64)
    c32 = COracle(*a, dtype=np.float32)
    for _ in range(10):
        c64.step(DT)
        c32.step(DT)
        assert circ_err(c32.theta, c64.theta).max() / TWO_PI < 1e-5
        assert (np.abs(c32.w - c64.w) / np.maximum(np.abs(c64.w), 1e-3)).max() < 1e-5


def test_oracle_rejects_delay_beyond_max():
    with pytest.raises(ValueError):
        tiny("numpy", 2, [(0, 1, 6, 0.5)], [0.0, 0.0], max_delay=5)


def test_ridge_fp64_matches_sklearn_and_exact_fit():
    """The FP64 ridge oracle against sklearn's Ridge(fit_intercept=True) (a library oracle,
    the estimator python/reservoir.py:81-82 uses), and exact recovery of a noiseless linear map."""
    from sklearn.linear_model import Ridge

    rng = np.random.default_rng(3)
    X = rng.standard_normal((400, 30))
    y = X @ rng.standard_normal(30) + 0.7 + 0.1 * rng.standard_normal(400)
    for alpha in (1e-4, 1.0, 50.0):
        w, c = oracle.ridge_fp64(X, y, alpha)
        ref = Ridge(alpha=alpha, fit_intercept=True, solver="cholesky").fit(X, y)
        assert np.abs(w - ref.coef_).max() < 1e-9
        assert abs(c - ref.intercept_) < 1e-9
    wt = rng.standard_normal(30)
    w, c = oracle.ridge_fp64(X, X @ wt - 2.5, 1e-12)
    assert np.abs(w - wt).max() < 1e-8 and abs(c + 2.5) < 1e-8


def test_bf16_round_known_values():
    """BF16 rounding: exact values stay, halfway cases go to even, relative error <= 2^-8 (half an ulp)."""
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 2**-7, 1.0 + 3 * 2**-8, -2.5, 0.0], np.float32)
    want = np.array([1.0, 1.0, 1.0 + 2**-7, 1.0 + 2**-6, -2.5, 0.0], np.float32)
    assert np.array_equal(oracle.bf16_round(x), want)
    r = np.random.default_rng(0).standard_normal(10000).astype(np.float32)
    assert (np.abs(oracle.bf16_round(r) - r) / np.abs(r)).max() <= 2**-8
