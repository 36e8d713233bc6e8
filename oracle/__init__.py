"""CPU oracle for LEVIATHAN's discrete step.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` import this package, and only as the
checker (or the timed CPU baseline), never as the computation being measured or
shipped.  The product library (``leviathan_v1_amd``) never imports it.

Three restatements of the reference step (src/cuda/leviathan_kernel.cu:391-438):

* ``NumpyOracle``  -- FP64, vectorised NumPy (this file).
* ``COracle(dtype=np.float64)`` -- FP64 C, sequential (leviathan_oracle.c).
* ``COracle(dtype=np.float64, parallel=True)`` -- the same FP64 C with OpenMP
  over rows and nodes; theta, W and g bit-identical to the sequential one, r
  summed in a fixed blocked order (for the 10^7-node workload).
* ``COracle(dtype=np.float32)`` -- FP32 C/OpenMP port of the reference's four
  kernels: the CPU baseline.

The two FP64 versions are written independently (gather by fancy indexing and
``np.bincount`` here, scalar loops in C) and are checked against each other and
against closed forms in ``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

TWO_PI = 2.0 * np.pi


@dataclass
class Params:
    """Physics parameters; defaults are leviathan_kernel.cu:281-297 (and the
    literal g_rest = 1.5 of :433).  MODEL.md:71-86 lists the same values."""

    eta: float = 0.01
    lam: float = 0.001
    gamma: float = 0.05
    alpha_eps: float = 0.1
    kappa_w: float = 0.01
    w_min: float = 0.0
    w_max: float = 2.0
    g0: float = 1.5
    r0: float = 0.5
    beta: float = 0.01
    mu: float = 0.01
    g_rest: float = 1.5
    g_min: float = 0.1
    g_max: float = 5.0
    gain_controller: bool = True


class _CParams(ctypes.Structure):
    _fields_ = [
        ("eta", ctypes.c_double),
        ("lam", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("alpha_eps", ctypes.c_double),
        ("kappa_w", ctypes.c_double),
        ("w_min", ctypes.c_double),
        ("w_max", ctypes.c_double),
        ("r0", ctypes.c_double),
        ("beta", ctypes.c_double),
        ("mu", ctypes.c_double),
        ("g_rest", ctypes.c_double),
        ("g_min", ctypes.c_double),
        ("g_max", ctypes.c_double),
        ("gain_controller", ctypes.c_int32),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(LIB_PATH) or _stale():
        subprocess.run(["make", "-s", "-B", "-C", HERE, "liboracle.so"], check=True)
    return LIB_PATH


def _stale() -> bool:
    so = os.path.getmtime(LIB_PATH)
    srcs = [os.path.join(HERE, f) for f in ("leviathan_oracle.c", "oracle_step.inc")]
    return any(os.path.getmtime(s) > so for s in srcs)


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


class _StateMixin:
    def _init_state(self, N, max_delay, row_ptr, col_idx, delays, weights, theta, theta_hat, omega, params, dtype):
        self.dtype = np.dtype(dtype)
        self.N = int(N)
        self.max_delay = int(max_delay)
        self.B = self.max_delay + 1  # leviathan_kernel.cu:277
        self.params = params or Params()
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
        self.col_idx = np.ascontiguousarray(col_idx, dtype=np.int32)
        self.delays = np.ascontiguousarray(delays, dtype=np.uint8)
        self.w = np.array(weights, dtype=self.dtype)
        self.theta = np.array(theta, dtype=self.dtype)
        self.theta_hat = np.array(theta_hat, dtype=self.dtype)
        self.omega = np.array(omega, dtype=self.dtype)
        self.omega_hat = np.array(omega, dtype=self.dtype)  # omega_hat0 = omega (:350)
        self.eps = np.zeros(self.N, dtype=self.dtype)
        self.S = np.zeros(self.N, dtype=self.dtype)
        self.buf = np.repeat(self.theta[:, None], self.B, axis=1).copy()  # [N][B] (:79-90)
        self.buf_idx = 0
        self.g = self.dtype.type(self.params.g0)
        self.gain_controller = bool(self.params.gain_controller)
        self.step_count = 0
        if self.row_ptr.shape != (self.N + 1,):
            raise ValueError("row_ptr must have N+1 entries")
        E = int(self.row_ptr[-1])
        if self.col_idx.shape != (E,) or self.delays.shape != (E,) or self.w.shape != (E,):
            raise ValueError("col_idx, delays and weights must have row_ptr[N] entries")
        if E and int(self.delays.max()) > self.max_delay:
            raise ValueError("a delay exceeds max_delay")

    # accessors mirroring leviathan_kernel.cu:440-497
    def get_theta(self):
        return self.theta.copy()

    def set_theta(self, theta):
        theta = np.asarray(theta, dtype=self.dtype)
        if theta.shape != (self.N,):
            raise RuntimeError("theta array size mismatch")
        self.theta[:] = theta
        self.buf[:] = self.theta[:, None]  # :445-453 re-fills the whole ring

    def set_gain(self, g):
        self.g = self.dtype.type(g)

    def get_gain(self):
        return float(self.g)

    def set_gain_controller(self, enabled):
        self.gain_controller = bool(enabled)

    def reset_weights(self, w):
        w = np.asarray(w, dtype=self.dtype)
        if w.shape != self.w.shape:
            raise RuntimeError("weights size mismatch")
        self.w[:] = w

    def history(self, delay):
        """theta(t - delay) for every node, as stored in the ring."""
        return self.buf[:, (self.buf_idx - delay) % self.B].copy()


class NumpyOracle(_StateMixin):
    """FP64 vectorised restatement of one reference step.

    prep    leviathan_kernel.cu:100-110
    physics leviathan_kernel.cu:124-166 (per-row sums via np.bincount)
    integ.  leviathan_kernel.cu:182-196, r :260-263, gain :431-435
    """

    def __init__(self, N, max_delay, row_ptr, col_idx, delays, weights, theta, theta_hat, omega, params=None):
        self._init_state(N, max_delay, row_ptr, col_idx, delays, weights, theta, theta_hat, omega, params, np.float64)
        self.rows = np.repeat(np.arange(self.N, dtype=np.int64), np.diff(self.row_ptr))
        self.drive_b = None

    def step(self, dt, u=0.0):
        p = self.params
        dt = float(dt)
        diff = self.theta - self.theta_hat
        diff = diff - np.floor(diff / TWO_PI + 0.5) * TWO_PI
        eps = np.sin(diff)
        self.theta_hat += dt * self.omega_hat
        self.omega_hat += p.kappa_w * dt * eps
        self.eps = eps

        rows = self.rows
        slot = (self.buf_idx - self.delays.astype(np.int64) + self.B) % self.B
        h = self.buf[self.col_idx, slot]
        d = h - self.theta[rows]
        d = d - np.floor(d / TWO_PI + 0.5) * TWO_PI
        s, c = np.sin(d), np.cos(d)
        w = self.w
        with np.errstate(invalid="ignore"):
            S = np.bincount(rows, weights=w * s, minlength=self.N)
            dw = dt * (p.eta * c * (1.0 - p.gamma * np.abs(eps[rows])) - p.lam * w)
            wn = w + dw
        wn[~np.isfinite(wn)] = p.w_min
        np.clip(wn, p.w_min, p.w_max, out=wn)
        self.w = wn
        self.S = S

        om = self.omega if self.drive_b is None else self.omega + self.drive_b * u
        with np.errstate(invalid="ignore"):
            x = self.theta + dt * (om + self.g * S - p.alpha_eps * eps)
            x = x - np.floor(x / TWO_PI) * TWO_PI
        x[~np.isfinite(x)] = 0.0
        nxt = (self.buf_idx + 1) % self.B
        self.theta = x
        self.buf[:, nxt] = x
        self.buf_idx = nxt
        r = float(np.hypot(np.cos(x).sum(), np.sin(x).sum()) / self.N)
        if self.gain_controller:
            g = self.g + dt * (p.beta * (p.r0 - r) - p.mu * (self.g - p.g_rest))
            self.g = float(min(max(g, p.g_min), p.g_max))
        self.step_count += 1
        return r

    def run(self, dt, n, u_seq=None):
        out = np.empty(n)
        for t in range(n):
            out[t] = self.step(dt, 0.0 if u_seq is None else u_seq[t])
        return out


class COracle(_StateMixin):
    """The C restatement (leviathan_oracle.c) with state held in NumPy arrays.

    dtype=np.float64: FP64 oracle.  dtype=np.float32: FP32 OpenMP port of the
    reference's kernels, used as the CPU baseline.
    """

    def __init__(self, N, max_delay, row_ptr, col_idx, delays, weights, theta, theta_hat, omega,
                 params=None, dtype=np.float64, parallel=False):
        self._init_state(N, max_delay, row_ptr, col_idx, delays, weights, theta, theta_hat, omega, params, dtype)
        lib = _load()
        sfx = ("f64p" if parallel else "f64") if self.dtype == np.float64 else "f32"
        self._run = getattr(lib, "lev_oracle_run_" + sfx)
        self._run.restype = None
        self._real = ctypes.c_double if self.dtype == np.float64 else ctypes.c_float
        self._bi = ctypes.c_int32(0)
        self._g = np.array([self.g], dtype=self.dtype)
        self.drive_b = None

    def _cparams(self):
        p = self.params
        return _CParams(p.eta, p.lam, p.gamma, p.alpha_eps, p.kappa_w, p.w_min, p.w_max, p.r0, p.beta,
                        p.mu, p.g_rest, p.g_min, p.g_max, int(self.gain_controller))

    def run(self, dt, n, u_seq=None):
        n = int(n)
        r_out = np.empty(n, dtype=self.dtype)
        cp = self._cparams()
        self._bi.value = self.buf_idx
        self._g[0] = self.g
        drive = None if self.drive_b is None else np.ascontiguousarray(self.drive_b, dtype=self.dtype)
        useq = None if u_seq is None else np.ascontiguousarray(u_seq[:n], dtype=self.dtype)
        self._run(ctypes.byref(cp), ctypes.c_int32(self.N), ctypes.c_int32(self.B),
                  _ptr(self.row_ptr), _ptr(self.col_idx), _ptr(self.delays), _ptr(self.w),
                  _ptr(self.theta), _ptr(self.theta_hat), _ptr(self.omega), _ptr(self.omega_hat),
                  _ptr(self.eps), _ptr(self.S), _ptr(self.buf), ctypes.byref(self._bi),
                  _ptr(self._g), self._real(dt), ctypes.c_int32(n), _ptr(r_out), _ptr(drive), _ptr(useq))
        self.buf_idx = self._bi.value
        self.g = self.dtype.type(self._g[0])
        self.step_count += n
        return r_out

    def step(self, dt, u=0.0):
        return float(self.run(dt, 1, None if self.drive_b is None else np.array([u]))[0])


def set_threads(n: int) -> int:
    """Threads of the OpenMP restatements from now on; returns the previous count."""
    lib = _load()
    prev = int(lib.lev_oracle_max_threads())
    lib.lev_oracle_set_threads(ctypes.c_int(int(n)))
    return prev


def circ_err(a, b):
    """|a - b| on the circle, in radians."""
    d = np.asarray(a, dtype=np.float64) - np.asarray(b, dtype=np.float64)
    return np.abs(d - np.round(d / TWO_PI) * TWO_PI)


def ridge_fp64(X, y, alpha):
    """FP64 ridge regression with an intercept (the estimator of sklearn's Ridge with
    fit_intercept=True, which python/reservoir.py:81-82 uses): centre X and y, solve
    (Xc^T Xc + alpha I) w = Xc^T yc by Cholesky, intercept = mean(y) - mean(X) w."""
    import scipy.linalg as sla

    X = np.asarray(X, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    mu = X.mean(axis=0)
    Xc = X - mu
    yc = y - y.mean()
    G = Xc.T @ Xc
    G[np.diag_indices_from(G)] += alpha
    w = sla.cho_solve(sla.cho_factor(G, lower=True), Xc.T @ yc)
    return w, float(y.mean() - mu @ w)


def bf16_round(x):
    """Round float32 values to the nearest BF16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32)
