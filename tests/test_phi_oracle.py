"""Pins the hub-statistics oracle (oracle/phi.py) before it checks the device.

The reference ships no tests or fixtures for this path, so each function is pinned by
what fixes it independently of its own code:

* binarize: hand-worked phases at the FP32 thresholds of leviathan_phi_kernel.cu:69-79,
  and, on [0, 2pi), the threshold python/leviathan_iit_integration.py:87 uses
  (phase > numpy.pi in FP64) -- a second definition that must give the same bits;
* counts: a histogram of consecutive state pairs by np.add.at (no loop over calls);
* state-by-node TPM: the permutation TPM of a cyclic state sequence (closed form),
  0.5 for states never left, and the reference's own FP64 summation order;
* phi: exact zero for the uniform TPM (all factors 1/2), a few FP64 ulps otherwise,
  and a planted mistake (a factor left out of p_B) that makes it large.
"""
import numpy as np
import pytest

from oracle import phi as P

F = np.float32


@pytest.mark.parametrize("theta, bit", [
    (F(np.pi), 1),                      # 3.14159274 > 3.14159250
    (F(3.1415926), 0),                  # the threshold itself
    (np.nextafter(F(3.1415926), F(0)), 0),
    (F(0.0), 0),
    (F(-0.0), 0),
    (np.nextafter(F(6.2831853), F(0)), 1),
    (F(6.2831853), 0),                  # fmod gives 0
    (F(-0.1), 1),                       # fmod -0.1, then + 2pi
    (F(-4.0), 0),                       # -4 + 2pi = 2.28
    (F(7.0), 0),                        # 7 - 2pi = 0.717
    (F(10.0), 1),                       # 10 - 2pi = 3.717
    (F(np.nan), 0),
])
def test_binarize_hand_points(theta, bit):
    assert P.binarize(np.array([theta], F), [0]) == bit


def test_binarize_matches_python_threshold_on_wrapped_phases():
    rng = np.random.default_rng(0)
    th = rng.uniform(0, 2 * np.pi, 200_000).astype(F)
    th = th[th < F(6.2831853)]
    # every float near the threshold too
    near = np.array([np.nextafter(F(np.pi), F(k)) for k in (-10, 10)] + [F(np.pi)], F)
    th = np.concatenate([th, near])
    bits_py = th.astype(np.float64) > np.pi
    bits_cu = np.array([P.binarize(th[i:i + 1], [0]) for i in range(0, th.size, 97)], bool)
    assert np.array_equal(bits_cu, bits_py[::97])
    # bit order: hub i is bit i
    assert P.binarize(np.array([4.0, 1.0, 4.0], F), [0, 1, 2]) == 0b101
    assert P.binarize(np.array([4.0, 1.0, 4.0], F), [1, 2]) == 0b10


def _states(thetas, hubs):
    x = np.fmod(thetas[:, hubs], P.TWO_PI_F)
    x = np.where(x < 0, (x + P.TWO_PI_F).astype(F), x)
    bits = (x > P.PI_F).astype(np.int64)
    return (bits << np.arange(len(hubs))).sum(axis=1)


def test_counts_are_the_pair_histogram():
    rng = np.random.default_rng(1)
    hubs = [3, 0, 7, 5]
    T = 3000
    thetas = rng.uniform(-3, 9, (T, 8)).astype(F)
    o = P.PhiOracle(hubs)
    seq = [o.accumulate(thetas[t]) for t in range(T)]
    st = _states(thetas, hubs)
    assert np.array_equal(seq, st)
    ref = np.zeros((16, 16), np.int64)
    np.add.at(ref, (st[:-1], st[1:]), 1)
    assert np.array_equal(o.counts, ref)
    assert np.array_equal(o.row_totals, ref.sum(axis=1))
    assert o.calls == T and o.prev == st[-1]
    o.reset()
    assert o.counts.sum() == 0 and o.prev == -1 and o.calls == 0


def test_sbn_cycle_closed_form_and_empty_rows():
    H = 5
    S = 1 << H
    counts = np.zeros((S, S), np.int64)
    for s in range(S - 3):  # states S-3 .. S-1 are never left
        counts[s, (s + 1) % S] = 3
    tot = counts.sum(axis=1)
    sbn = P.sbn_tpm(counts, tot, H)
    for s in range(S):
        for n in range(H):
            want = ((s + 1) >> n) & 1 if s < S - 3 else 0.5
            assert sbn[s, n] == want


def test_sbn_exact_quotient_vs_reference_summation_order():
    rng = np.random.default_rng(2)
    H = 6
    S = 1 << H
    counts = rng.integers(0, 50, (S, S)).astype(np.int64)
    counts[rng.random(S) < 0.2] = 0
    tot = counts.sum(axis=1)
    a = P.sbn_tpm(counts, tot, H)
    b = P.sbn_tpm_reference_order(counts, tot, H)
    assert np.abs(a - b).max() < 1e-14
    num = P.sbn_numerators(counts, H)
    nz = tot > 0
    # the exact rational value num / tot, rounded once
    assert np.array_equal(a[nz], num[nz] / tot[nz, None])


def test_phi_uniform_tpm_is_exactly_zero():
    for H in (2, 3, 6, 10):
        S = 1 << H
        sbn = P.sbn_tpm(np.zeros((S, S), np.int64), np.zeros(S, np.int64), H)
        part = P.phi_partitions(sbn, H)
        assert part.shape == ((1 << (H - 1)) - 1,)
        assert np.all(part == 0.0)
        assert P.find_mip(part) == (0.0, 1)


def test_phi_is_rounding_residue_and_a_dropped_factor_is_not():
    rng = np.random.default_rng(3)
    H = 7
    S = 1 << H
    counts = rng.integers(0, 20, (S, S)).astype(np.int64)
    sbn = P.sbn_tpm(counts, counts.sum(axis=1), H)
    part = P.phi_partitions(sbn, H)
    assert np.all(part >= 0) and part.max() < 1e-13
    # planted mistake: hub 0's factor left out of p_B when hub 0 is in part B
    states = np.arange(S)
    mask = 2  # hub 1 in A, hub 0 in B
    pi = np.ones(S)
    pa = np.ones(S)
    pb = np.ones(S)
    for n in range(H):
        prob = np.maximum(np.where((states >> n) & 1, sbn[:, n], 1 - sbn[:, n]), P.PROB_FLOOR)
        pi *= prob
        if (mask >> n) & 1:
            pa *= prob
        elif n != 0:
            pb *= prob
    bad = np.abs(np.log(np.maximum(pi, P.LOG_FLOOR)) - np.log(np.maximum(pa * pb, P.LOG_FLOOR))).mean()
    assert bad > 0.1 and abs(part[mask - 1]) < 1e-13


def test_single_hub_has_no_bipartition():
    o = P.PhiOracle([4])
    assert o.P == 0
    for v in (1.0, 4.0, 4.0, 1.0):
        o.accumulate(np.full(5, v, F))
    phi, mask, part = o.compute()
    assert (phi, mask) == (P.INVALID, 0) and part.size == 0
    assert o.counts.tolist() == [[0, 1], [1, 1]]


def test_limits():
    with pytest.raises(ValueError):
        P.PhiOracle([])
    with pytest.raises(ValueError):
        P.PhiOracle(list(range(11)))
    assert P.PhiOracle(list(range(10))).S == 1024
