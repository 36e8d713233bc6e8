"""Reservoir computing on the GPU-resident network.

Counterpart of python/reservoir.py:21-94 (``ReservoirComputer``), redesigned:

* input enters as a frequency drive omega_i(t) = omega_i + b_i u(t) on the
  device (the reference kicks phases through get/set_theta, which copies all
  phases over PCIe twice per step and erases the delay history each time,
  python/reservoir.py:45-53);
* features are cos/sin of the chosen nodes, captured on the device after
  every step (the reference copies all N phases back, :55-60);
* the ridge readout (sklearn Ridge with an intercept, :81-82) is solved from
  the normal equations: the Gram matrix of the centred features is formed
  on the matrix cores from BF16 inputs with FP32 accumulation, factorised
  once in FP64 on the host, and used to precondition a conjugate-gradient
  refinement whose residuals b - (Xc^T Xc + alpha I) w are computed on the
  device in FP64 from the FP32 features.  The result is the FP64 ridge
  solution; the BF16 Gram only sets how fast the refinement converges.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from .engine import LeviathanEngine
from .workloads import Workload


def multisine_input(T: int, n_sines: int = 5, period_lo: float = 20.0, period_hi: float = 200.0,
                    noise: float = 0.05, seed: int = 0) -> np.ndarray:
    """u(t) = (1/n) sum_k sin(2 pi t / P_k + phi_k) + N(0, noise^2), P_k ~ U[period_lo, period_hi] steps."""
    rng = np.random.Generator(np.random.PCG64(seed))
    P = rng.uniform(period_lo, period_hi, n_sines)
    phi = rng.uniform(0.0, 2 * np.pi, n_sines)
    t = np.arange(T, dtype=np.float64)[:, None]
    u = np.sin(2 * np.pi * t / P + phi).mean(axis=1) + rng.normal(0.0, noise, T)
    return u.astype(np.float32)


@dataclass
class RidgeResult:
    w: np.ndarray            # feature weights (FP64)
    intercept: float
    iterations: int
    rel_residual: float
    shift: float             # diagonal shift that made the BF16 Gram factorisable
    timings: dict = field(default_factory=dict)


class LeviathanReservoir:
    """A driven network whose node phases feed a linear readout."""

    def __init__(self, workload: Workload, feature_nodes, drive_b, max_rows: int, **engine_kwargs):
        self.wl = workload
        self.engine = LeviathanEngine(*workload.engine_args(), **engine_kwargs)
        self.engine.set_drive(drive_b)
        self.engine.features_setup(feature_nodes, max_rows)
        self.F = 2 * len(feature_nodes)

    def run(self, u) -> np.ndarray:
        """One driven step per input sample; appends a feature row per step; returns r."""
        return self.engine.step_n_features(self.wl.dt, u=u)

    @property
    def rows(self) -> int:
        return self.engine.feature_rows

    def features(self, row0: int, nrows: int) -> np.ndarray:
        return self.engine.features_get(row0, nrows)

    # -------------------------------------------------------------- readout
    def _A(self, row0, nrows, mean, alpha, x):
        v = self.engine.features_matvec(row0, nrows, mean, x)           # Xc x
        g = self.engine.features_colsum(row0, nrows, v) - mean * v.sum()  # Xc^T v
        return g + alpha * x

    def _factor(self, row0, nrows, mean, alpha, terms):
        """Cholesky factor, on the device, of the matrix-core Gram + alpha I with the smallest diagonal
        shift that makes it factorisable; returns (shift, Gram kernel ms)."""
        return self.engine.features_factor(row0, nrows, mean, alpha, terms=terms)

    def _pcg(self, b, x, row0, nrows, mean, alpha, tol, max_iter, stall):
        """Preconditioned conjugate gradients on the exact FP64 operator from x, preconditioned by the
        device factor; stops at tol, at max_iter, or when the residual fell by less than 2x over the
        last ``stall`` iterations."""
        bn = float(np.linalg.norm(b)) or 1.0
        r = b - self._A(row0, nrows, mean, alpha, x)
        rel = float(np.linalg.norm(r)) / bn
        hist = [rel]
        if rel <= tol:
            return x, 0, rel
        z = self.engine.features_precond(r)
        p = z.copy()
        rz = float(r @ z)
        it = 0
        while it < max_iter:
            Ap = self._A(row0, nrows, mean, alpha, p)
            a = rz / float(p @ Ap)
            x = x + a * p
            r -= a * Ap
            it += 1
            rel = float(np.linalg.norm(r)) / bn
            hist.append(rel)
            if rel <= tol or (stall and it >= stall and rel > 0.5 * hist[-1 - stall]):
                break
            z = self.engine.features_precond(r)
            rz_new = float(r @ z)
            p = z + (rz_new / rz) * p
            rz = rz_new
        return x, it, rel

    def fit(self, target, row0: int, alpha: float = 1e-4, tol: float = 1e-10, max_iter: int = 500,
            gram_terms=None) -> RidgeResult:
        """Ridge regression with intercept on rows [row0, row0 + len(target)).

        gram_terms: 1 (BF16 Gram), 3 (split hi/lo BF16 Gram), 0 (FP64 Gram) or None: start from the
        BF16 Gram and switch to the split Gram, then to the FP64 Gram, when the refinement stalls
        (ill-conditioned features)."""
        y = np.asarray(target, dtype=np.float64)
        nrows = y.shape[0]
        t0 = time.perf_counter()
        mean = self.engine.features_colsum(row0, nrows) / nrows
        ybar = float(y.mean())
        yc = y - ybar
        b = self.engine.features_colsum(row0, nrows, yc) - mean * yc.sum()
        t1 = time.perf_counter()
        plan = [gram_terms] if gram_terms is not None else [1, 3, 0]
        timings = {"moments_s": t1 - t0, "factor_s": 0.0, "refine_s": 0.0}
        x = None
        its = 0
        shift = 0.0
        for k, terms in enumerate(plan):
            t2 = time.perf_counter()
            shift, gram_ms = self._factor(row0, nrows, mean, alpha, terms)
            t3 = time.perf_counter()
            if x is None:
                x = self.engine.features_precond(b)
            last = k == len(plan) - 1
            x, it, rel = self._pcg(b, x, row0, nrows, mean, alpha, tol, max_iter - its, 0 if last else 25)
            its += it
            t4 = time.perf_counter()
            timings["factor_s"] += t3 - t2  # Gram on the matrix cores + FP64 Cholesky on the device
            timings["refine_s"] += t4 - t3
            timings.setdefault("gram_ms_by_terms", {})[terms] = gram_ms
            timings["gram_terms"] = terms
            if rel <= tol:
                break
        # true residual of the final iterate; the recursive residual drifts from it on ill-conditioned
        # features, so restart from the iterate (which recomputes the residual exactly) while it is
        # above tol and iterations remain
        rel = float(np.linalg.norm(b - self._A(row0, nrows, mean, alpha, x))) / (float(np.linalg.norm(b)) or 1.0)
        restarts = 0
        while rel > tol and its < max_iter and restarts < 3:
            t3 = time.perf_counter()
            x, it, _ = self._pcg(b, x, row0, nrows, mean, alpha, tol, max_iter - its, 0)
            its += it
            restarts += 1
            timings["refine_s"] += time.perf_counter() - t3
            rel = float(np.linalg.norm(b - self._A(row0, nrows, mean, alpha, x))) / (float(np.linalg.norm(b)) or 1.0)
        timings["restarts"] = restarts
        intercept = ybar - float(mean @ x)
        return RidgeResult(x, intercept, its, rel, shift, timings)

    def predict(self, res: RidgeResult, row0: int, nrows: int) -> np.ndarray:
        zero = np.zeros(self.F)
        return self.engine.features_matvec(row0, nrows, zero, res.w) + res.intercept

    def close(self):
        self.engine.close()


def nrmse(pred, y) -> float:
    y = np.asarray(y, dtype=np.float64)
    return float(np.sqrt(np.mean((np.asarray(pred, np.float64) - y) ** 2)) / (np.std(y) or 1.0))
