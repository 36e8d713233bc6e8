"""CPU oracle for the hub transition statistics and the bipartition measure.

TEST INFRASTRUCTURE ONLY, like the rest of this package: only ``tests/`` import
it, as the checker of ``libleviathan_hip.so``'s ``phi_*`` entry points.

Restates src/cuda/leviathan_phi_kernel.cu:

* ``binarize``     -- :69-79   bit i of the state is set when fmodf(theta, 6.2831853f),
                              moved into [0, 2pi), exceeds 3.1415926f (FP32, exact).
* ``PhiOracle.accumulate`` -- :81-91 one transition prev -> state per call after
                              the first; :290-299 counts calls.
* ``sbn_tpm``      -- :98-136  P(hub n on at t+1 | state s at t); a state never left
                              gets the uniform row 1/S, whose column sums are 0.5.
                              Also python/leviathan_iit_integration.py:114-149 (the
                              same formula in FP64).
* ``phi_partitions`` -- :143-207 per bipartition mask = p + 1 (p < 2^(H-1) - 1):
                              sum over states of |log p_int - log(p_A p_B)| / S, with
                              every factor clamped at 1e-10 and both products at 1e-30.
* ``find_mip``     -- :213-235 the minimum over the bipartitions (1e30 when there are none).

Pinned in tests/test_phi_oracle.py against hand-worked states at the FP32 thresholds,
np.add.at histograms of state sequences, the permutation TPM of a cyclic sequence
(closed form), the uniform rows of an empty TPM and the exact zero that p_A p_B = p_int
forces when both products are formed in one order.
"""
from __future__ import annotations

import numpy as np

MAX_HUBS = 10                      # leviathan_phi_kernel.cu:243-246
TWO_PI_F = np.float32(6.2831853)   # :74 (rounds to the float 6.28318548)
PI_F = np.float32(3.1415926)       # :77 (rounds to the float 3.14159250)
PROB_FLOOR = 1e-10                 # :189
LOG_FLOOR = 1e-30                  # :201-202
INVALID = 1e30                     # :171 and :220


def binarize(theta, hubs) -> int:
    """Hub state of one phase vector (leviathan_phi_kernel.cu:69-79), FP32 throughout."""
    th = np.asarray(theta, dtype=np.float32)
    state = 0
    for i, node in enumerate(hubs):
        x = np.fmod(th[int(node)], TWO_PI_F)
        if x < 0:
            x = np.float32(x + TWO_PI_F)
        if x > PI_F:
            state |= 1 << i
    return state


class PhiOracle:
    """Host restatement of one PhiHandle (leviathan_phi_kernel.cu:27-53, :242-352)."""

    def __init__(self, hubs):
        self.hubs = [int(h) for h in hubs]
        H = len(self.hubs)
        if not 1 <= H <= MAX_HUBS:
            raise ValueError(f"1 <= num_hubs <= {MAX_HUBS} required, got {H}")
        self.H = H
        self.S = 1 << H
        self.P = (1 << (H - 1)) - 1
        self.reset()

    def reset(self):
        """:340-348."""
        self.counts = np.zeros((self.S, self.S), dtype=np.int64)
        self.row_totals = np.zeros(self.S, dtype=np.int64)
        self.prev = -1
        self.calls = 0

    def accumulate(self, theta) -> int:
        """:81-91 and :290-299: record prev -> state (after the first call) and count the call."""
        state = binarize(theta, self.hubs)
        if 0 <= self.prev < self.S:
            self.counts[self.prev, state] += 1
            self.row_totals[self.prev] += 1
        self.prev = state
        self.calls += 1
        return state

    def sbn_tpm(self) -> np.ndarray:
        return sbn_tpm(self.counts, self.row_totals, self.H)

    def compute(self):
        """(phi of the minimum-information bipartition, its mask, per-partition values)."""
        part = phi_partitions(self.sbn_tpm(), self.H)
        phi, mask = find_mip(part)
        return phi, mask, part


def sbn_numerators(counts, H) -> np.ndarray:
    """[S, H] integer sums of counts[s, t] over the targets t with bit n set."""
    S = counts.shape[0]
    bits = (np.arange(S)[:, None] >> np.arange(H)[None, :]) & 1
    return counts.astype(np.int64) @ bits.astype(np.int64)


def sbn_tpm(counts, row_totals, H) -> np.ndarray:
    """State-by-node TPM [S, H] (leviathan_phi_kernel.cu:98-136): each entry is the
    exactly rounded quotient of the integer numerator by the row total; an empty row
    gives 0.5 (the uniform row 1/S summed over the S/2 targets with the bit set)."""
    num = sbn_numerators(counts, H).astype(np.float64)
    tot = np.asarray(row_totals, dtype=np.int64)
    out = np.full(num.shape, 0.5)
    nz = tot > 0
    out[nz] = num[nz] / tot[nz, None].astype(np.float64)
    return out


def sbn_tpm_reference_order(counts, row_totals, H) -> np.ndarray:
    """The same quantity summed as python/leviathan_iit_integration.py:121-147 does:
    per-target quotients count / total added in target order (FP64)."""
    S = counts.shape[0]
    full = np.empty((S, S))
    for s in range(S):
        tot = int(row_totals[s])
        full[s] = counts[s] / tot if tot > 0 else 1.0 / S
    out = np.zeros((S, H))
    for n in range(H):
        on = ((np.arange(S) >> n) & 1).astype(np.float64)
        for t in range(S):
            out[:, n] += full[:, t] * on[t]
    return out


def phi_partitions(sbn, H) -> np.ndarray:
    """phi of every bipartition mask = p + 1, p < 2^(H-1) - 1 (leviathan_phi_kernel.cu:143-207),
    in FP64: factors multiplied in hub order, into p_int and into p_A or p_B."""
    S = 1 << H
    P = (1 << (H - 1)) - 1
    sbn = np.asarray(sbn, dtype=np.float64).reshape(S, H)
    states = np.arange(S)
    out = np.empty(P)
    for p in range(P):
        mask = p + 1
        nA = bin(mask & (S - 1)).count("1")
        if nA == 0 or nA == H:
            out[p] = INVALID
            continue
        pi = np.ones(S)
        pa = np.ones(S)
        pb = np.ones(S)
        for n in range(H):
            bit = (states >> n) & 1
            prob = np.where(bit == 1, sbn[:, n], 1.0 - sbn[:, n])
            prob = np.maximum(prob, PROB_FLOOR)
            pi = pi * prob
            if (mask >> n) & 1:
                pa = pa * prob
            else:
                pb = pb * prob
        pp = pa * pb
        d = np.abs(np.log(np.maximum(pi, LOG_FLOOR)) - np.log(np.maximum(pp, LOG_FLOOR)))
        out[p] = d.sum() / S
    return out


def find_mip(part):
    """(minimum, its mask = index + 1) over the bipartitions; (1e30, 0) when there are none."""
    part = np.asarray(part, dtype=np.float64)
    if part.size == 0:
        return INVALID, 0
    k = int(np.argmin(part))
    return float(part[k]), k + 1
