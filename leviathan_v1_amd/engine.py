"""``LeviathanEngine``: the Python surface of the gfx950 step library.

Mirrors the reference's pybind11 class (src/bindings/leviathan_pybind11.cpp:27-104):
same constructor arguments, same methods, same error behaviour (a
``RuntimeError`` for a wrong-sized ``set_theta``), with the GIL released in
every call (ctypes does this for CDLL functions).  Arrays are cast to the
expected dtype and made contiguous like ``py::array_t<T>``'s forcecast, and,
unlike the reference, their lengths are checked.

Extensions: keyword options (history format and layout, physics parameters,
ensembles), ``step_n``, ``get_weights``, ``get_predictor``, ``get_history``,
``profile_steps`` and an input drive for reservoir use.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _lib
from . import checkpoint as _checkpoint

_PARAM_NAMES = ("eta", "lam", "gamma", "alpha_eps", "kappa_w", "w_min", "w_max", "g0", "r0", "beta", "mu",
                "g_rest", "g_min", "g_max")


def _arr(a, dtype, n: Optional[int], name: str):
    a = np.ascontiguousarray(a, dtype=dtype)
    if a.ndim != 1:
        raise ValueError(f"{name} must be one-dimensional")
    if n is not None and a.shape[0] != n:
        raise ValueError(f"{name} has {a.shape[0]} entries, expected {n}")
    return a


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class LeviathanEngine:
    """GPU-resident delayed Kuramoto network with adaptive coupling.

    ``LeviathanEngine(N, max_delay, row_ptr, col_idx, delays, weights, theta,
    theta_hat, omega)`` -- exactly the reference constructor
    (leviathan_pybind11.cpp:32-42, :86-89).

    Keyword extensions:
      history: "f32" (default) or "u16" (exact-wrap uint16 angles)
      layout:  "auto" (default: "node" for the direct schedule on a graph without delay-0
               edges and with little index locality, where paired row passes read a node's
               two slots from one line; "slot" otherwise), "slot" (ring[slot][node]) or
               "node" (ring[node][slot])
      edges_per_thread: 4, 8 or 16 (stream-block capacity 256 * this)
      schedule: "direct" (one fused kernel, per-edge gather), "binned" (source-
               ordered scatter into destination buckets, then the fused row
               update), "staged" (source-ordered scatter into each row block's
               own range, staged through LDS by the row update) or "auto"
               (all built; autotune() picks)
      bucket_edges: binned schedule bucket capacity (0: automatic)
      graph_chunk: steps per captured HIP graph in step_n (0 disables graphs)
      params: dict of physics parameters (names of LeviathanConfig)
      n_nets, net_offsets, g0, controller: ensemble of independent networks
    """

    def __init__(self, N, max_delay, row_ptr, col_idx, delays, weights, theta, theta_hat, omega, *,
                 history: str = "f32", layout: str = "auto", edges_per_thread: int = 8, graph_chunk: int = 100,
                 schedule: str = "direct", bucket_edges: int = 0,
                 params: Optional[dict] = None, n_nets: int = 1, net_offsets: Optional[Sequence[int]] = None,
                 g0: Optional[Sequence[float]] = None, controller: Optional[Sequence[bool]] = None):
        self._h = None
        lib = _lib.lib()
        N = int(N)
        if N <= 0:
            raise ValueError("N must be positive")
        row_ptr = _arr(row_ptr, np.int32, N + 1, "row_ptr")
        E = int(row_ptr[-1])
        col_idx = _arr(col_idx, np.int32, E, "col_idx")
        delays = _arr(delays, np.uint8, E, "delays")
        weights = _arr(weights, np.float32, E, "weights")
        theta = _arr(theta, np.float32, N, "theta")
        theta_hat = _arr(theta_hat, np.float32, N, "theta_hat")
        omega = _arr(omega, np.float32, N, "omega")
        self.N, self.E, self.max_delay = N, E, int(max_delay)
        self.n_nets = int(n_nets)
        # identifies the graph in checkpoints (restore refuses another graph's state)
        self.graph_digest = _checkpoint.graph_digest(row_ptr, col_idx, delays)
        extended = (history != "f32" or layout != "auto" or edges_per_thread != 8 or graph_chunk != 100
                    or schedule != "direct" or bucket_edges
                    or params or n_nets != 1 or net_offsets is not None or g0 is not None
                    or controller is not None)
        if not extended:
            h = lib.leviathan_init(N, self.max_delay, _p(row_ptr), _p(col_idx), _p(delays), _p(weights),
                                   _p(theta), _p(theta_hat), _p(omega))
        else:
            cfg = _lib.default_config()
            if history not in _lib.HISTORY:
                raise ValueError(f"history must be one of {list(_lib.HISTORY)}")
            if layout not in _lib.LAYOUT:
                raise ValueError(f"layout must be one of {list(_lib.LAYOUT)}")
            cfg.history = _lib.HISTORY[history]
            cfg.layout = _lib.LAYOUT[layout]
            cfg.edges_per_thread = int(edges_per_thread)
            cfg.graph_chunk = int(graph_chunk)
            if schedule not in _lib.SCHEDULE:
                raise ValueError(f"schedule must be one of {list(_lib.SCHEDULE)}")
            cfg.schedule = _lib.SCHEDULE[schedule]
            cfg.bucket_edges = int(bucket_edges)
            for k, v in (params or {}).items():
                if k == "gain_controller":
                    cfg.gain_controller = int(bool(v))
                elif k in _PARAM_NAMES:
                    setattr(cfg, k, float(v))
                else:
                    raise ValueError(f"unknown parameter {k!r}")
            offs = None if net_offsets is None else _arr(net_offsets, np.int32, self.n_nets + 1, "net_offsets")
            g0a = None if g0 is None else _arr(g0, np.float32, self.n_nets, "g0")
            ctl = None if controller is None else _arr(np.asarray(controller, dtype=bool).astype(np.int32),
                                                       np.int32, self.n_nets, "controller")
            h = lib.leviathan_init_ex(ctypes.byref(cfg), N, self.max_delay, _p(row_ptr), _p(col_idx), _p(delays),
                                      _p(weights), _p(theta), _p(theta_hat), _p(omega), self.n_nets, _p(offs),
                                      _p(g0a), _p(ctl))
        if not h:
            raise RuntimeError("leviathan_init failed: " + _lib.last_error())
        self._h = ctypes.c_void_p(h)
        self.history = history
        # the layout the handle holds ("auto" resolved)
        self.layout = {v: k for k, v in _lib.LAYOUT.items()}[lib.leviathan_get_layout(self._h)]

    # ------------------------------------------------------------------ reference surface
    def step(self, dt: float) -> float:
        """One step; returns r (leviathan_pybind11.cpp:44-47)."""
        r = _lib.lib().leviathan_step(self._h, float(dt))
        if r != r:  # NaN signals failure
            _lib.check("step")
        return float(r)

    def get_theta(self) -> np.ndarray:
        """Copy of the phases as a new float32 array (leviathan_pybind11.cpp:50-55)."""
        out = np.empty(self.N, dtype=np.float32)
        _lib.lib().leviathan_get_theta(self._h, _p(out))
        _lib.check("get_theta")
        return out

    def set_theta(self, theta) -> None:
        """Upload phases and reset the history (leviathan_pybind11.cpp:57-65)."""
        theta = np.ascontiguousarray(theta, dtype=np.float32)
        if theta.size != self.N:
            raise RuntimeError(f"theta array size mismatch: expected {self.N}, got {theta.size}")
        _lib.lib().leviathan_set_theta(self._h, _p(theta.reshape(-1)))
        _lib.check("set_theta")

    def set_gain(self, g: float) -> None:
        _lib.lib().leviathan_set_gain(self._h, float(g))
        _lib.check("set_gain")

    def get_gain(self) -> float:
        g = _lib.lib().leviathan_get_gain(self._h)
        _lib.check("get_gain")
        return float(g)

    def set_gain_controller(self, enabled: bool) -> None:
        _lib.lib().leviathan_set_gain_controller(self._h, bool(enabled))
        _lib.check("set_gain_controller")

    def get_vram_usage(self) -> int:
        return int(_lib.lib().leviathan_get_vram_usage(self._h))

    def reset_weights(self, weights) -> None:
        weights = _arr(weights, np.float32, self.E, "weights")
        _lib.lib().leviathan_reset_weights(self._h, _p(weights))
        _lib.check("reset_weights")

    # ------------------------------------------------------------------ extensions
    def step_n(self, dt: float, n: int) -> np.ndarray:
        """n steps with one synchronisation; returns r as float32[n] (or [n, n_nets])."""
        n = int(n)
        out = np.empty((n, self.n_nets), dtype=np.float32)
        if _lib.lib().leviathan_step_n(self._h, float(dt), n, _p(out)) != 0:
            _lib.check("step_n")
        return out[:, 0] if self.n_nets == 1 else out

    def step_n_driven(self, dt: float, u) -> np.ndarray:
        u = _arr(u, np.float32, None, "u")
        n = u.shape[0]
        out = np.empty((n, self.n_nets), dtype=np.float32)
        if _lib.lib().leviathan_step_n_driven(self._h, float(dt), n, _p(u), _p(out)) != 0:
            _lib.check("step_n_driven")
        return out[:, 0] if self.n_nets == 1 else out

    def set_drive(self, b) -> None:
        """omega_i(t) = omega_i + b_i * u_t (None switches the drive off)."""
        arr = None if b is None else _arr(b, np.float32, self.N, "drive_b")
        if _lib.lib().leviathan_set_drive(self._h, _p(arr)) != 0:
            _lib.check("set_drive")

    def get_weights(self) -> np.ndarray:
        out = np.empty(self.E, dtype=np.float32)
        if _lib.lib().leviathan_get_weights(self._h, _p(out)) != 0:
            _lib.check("get_weights")
        return out

    def get_predictor(self):
        th = np.empty(self.N, dtype=np.float32)
        om = np.empty(self.N, dtype=np.float32)
        if _lib.lib().leviathan_get_predictor(self._h, _p(th), _p(om)) != 0:
            _lib.check("get_predictor")
        return th, om

    def get_history(self, delay: int) -> np.ndarray:
        out = np.empty(self.N, dtype=np.float32)
        if _lib.lib().leviathan_get_history(self._h, int(delay), _p(out)) != 0:
            _lib.check("get_history")
        return out

    # ------------------------------------------------------------------ full state
    def set_predictor(self, theta_hat, omega_hat) -> None:
        thh = _arr(theta_hat, np.float32, self.N, "theta_hat")
        omh = _arr(omega_hat, np.float32, self.N, "omega_hat")
        if _lib.lib().leviathan_set_predictor(self._h, _p(thh), _p(omh)) != 0:
            _lib.check("set_predictor")

    def set_history(self, delay: int, values) -> None:
        """theta(t - delay) as the next step will read it, for 1 <= delay <= max_delay."""
        v = _arr(values, np.float32, self.N, "values")
        if _lib.lib().leviathan_set_history(self._h, int(delay), _p(v)) != 0:
            _lib.check("set_history")

    def get_gains(self) -> np.ndarray:
        """Each network's gain in FP64 (the controller's own precision)."""
        out = np.empty(self.n_nets, dtype=np.float64)
        if _lib.lib().leviathan_get_gains(self._h, _p(out)) != 0:
            _lib.check("get_gains")
        return out

    def set_gains(self, g) -> None:
        g = _arr(g, np.float64, self.n_nets, "g")
        if _lib.lib().leviathan_set_gains(self._h, _p(g)) != 0:
            _lib.check("set_gains")

    def set_step_count(self, step: int) -> None:
        if _lib.lib().leviathan_set_step_count(self._h, int(step)) != 0:
            _lib.check("set_step_count")

    def checkpoint(self) -> dict:
        """Everything the next steps depend on: phases, predictor, weights, the history seen by
        delays 1..max_delay, FP64 gains and the step count (see leviathan_v1_amd.checkpoint)."""
        thh, omh = self.get_predictor()
        D = self.max_delay
        hist = np.empty((D, self.N), dtype=np.float32)
        for d in range(1, D + 1):
            hist[d - 1] = self.get_history(d)
        return {"theta": self.get_theta(), "theta_hat": thh, "omega_hat": omh, "weights": self.get_weights(),
                "history": hist, "gains": self.get_gains(), "step": np.int64(self.step_count),
                "N": np.int64(self.N), "E": np.int64(self.E), "max_delay": np.int64(D),
                "graph_digest": np.array(self.graph_digest)}

    def restore(self, state: dict) -> None:
        """Continue from a checkpoint() of an engine with the same graph and delays."""
        if (int(state["N"]), int(state["E"]), int(state["max_delay"])) != (self.N, self.E, self.max_delay):
            raise ValueError("checkpoint is of a different graph (N, E or max_delay differ)")
        if "graph_digest" in state and str(state["graph_digest"]) != self.graph_digest:
            raise ValueError("checkpoint was saved from a different graph (CSR or delays differ)")
        self.set_step_count(int(state["step"]))
        self.reset_weights(state["weights"])
        self.set_theta(state["theta"])
        self.set_predictor(state["theta_hat"], state["omega_hat"])
        hist = np.asarray(state["history"])
        for d in range(1, self.max_delay + 1):
            self.set_history(d, hist[d - 1])
        self.set_gains(state["gains"])

    # ------------------------------------------------------------------ reservoir features
    def features_setup(self, nodes, max_rows: int) -> None:
        nodes = _arr(nodes, np.int32, None, "nodes")
        self.n_feat_nodes = int(nodes.shape[0])
        if _lib.lib().leviathan_features_setup(self._h, _p(nodes), self.n_feat_nodes, int(max_rows)) != 0:
            _lib.check("features_setup")

    def step_n_features(self, dt: float, n: int = None, u=None) -> np.ndarray:
        """n steps (or len(u) driven steps) that each append one feature row."""
        ua = None if u is None else _arr(u, np.float32, None, "u")
        n = int(ua.shape[0]) if ua is not None else int(n)
        out = np.empty((n, self.n_nets), dtype=np.float32)
        if _lib.lib().leviathan_step_n_features(self._h, float(dt), n, _p(ua), _p(out)) != 0:
            _lib.check("step_n_features")
        return out[:, 0] if self.n_nets == 1 else out

    @property
    def feature_rows(self) -> int:
        return int(_lib.lib().leviathan_features_rows(self._h))

    def features_get(self, row0: int, nrows: int) -> np.ndarray:
        out = np.empty((int(nrows), 2 * self.n_feat_nodes), dtype=np.float32)
        if _lib.lib().leviathan_features_get(self._h, int(row0), int(nrows), _p(out)) != 0:
            _lib.check("features_get")
        return out

    def features_colsum(self, row0: int, nrows: int, weights=None) -> np.ndarray:
        w = None if weights is None else _arr(weights, np.float64, int(nrows), "weights")
        out = np.empty(2 * self.n_feat_nodes, dtype=np.float64)
        if _lib.lib().leviathan_features_colsum(self._h, int(row0), int(nrows), _p(w), _p(out)) != 0:
            _lib.check("features_colsum")
        return out

    def features_matvec(self, row0: int, nrows: int, mean, w) -> np.ndarray:
        F = 2 * self.n_feat_nodes
        mean = _arr(mean, np.float64, F, "mean")
        w = _arr(w, np.float64, F, "w")
        out = np.empty(int(nrows), dtype=np.float64)
        if _lib.lib().leviathan_features_matvec(self._h, int(row0), int(nrows), _p(mean), _p(w), _p(out)) != 0:
            _lib.check("features_matvec")
        return out

    def features_gram(self, row0: int, nrows: int, mean, terms: int = 1):
        """(G, ms): Gram of the centred features on the matrix cores (BF16 inputs, FP32
        accumulation; terms=3 uses split hi/lo BF16 inputs), and the Gram kernel's time."""
        F = 2 * self.n_feat_nodes
        mean = _arr(mean, np.float64, F, "mean")
        g = np.empty((F, F), dtype=np.float32)
        ms = ctypes.c_double(0.0)
        if _lib.lib().leviathan_features_gram(self._h, int(row0), int(nrows), _p(mean), int(terms), _p(g),
                                              ctypes.byref(ms)) != 0:
            _lib.check("features_gram")
        return g, ms.value

    def features_factor(self, row0: int, nrows: int, mean, alpha: float, terms: int = 1):
        """FP64 Cholesky factor of Gram + (alpha + shift) I on the device, kept for features_precond;
        terms 1 or 3: the matrix-core Gram (as features_gram), 0: the Gram formed in FP64.
        Returns (shift, Gram ms)."""
        F = 2 * self.n_feat_nodes
        mean = _arr(mean, np.float64, F, "mean")
        shift = ctypes.c_double(0.0)
        ms = ctypes.c_double(0.0)
        if _lib.lib().leviathan_features_factor(self._h, int(row0), int(nrows), _p(mean), int(terms), float(alpha),
                                                ctypes.byref(shift), ctypes.byref(ms)) != 0:
            _lib.check("features_factor")
        return shift.value, ms.value

    def features_precond(self, r) -> np.ndarray:
        """(Gram + (alpha + shift) I)^{-1} r with the factor of the last features_factor."""
        r = _arr(r, np.float64, 2 * self.n_feat_nodes, "r")
        z = np.empty_like(r)
        if _lib.lib().leviathan_features_precond(self._h, _p(r), _p(z)) != 0:
            _lib.check("features_precond")
        return z

    def set_gain_net(self, net: int, g: float) -> None:
        if _lib.lib().leviathan_set_gain_net(self._h, int(net), float(g)) != 0:
            _lib.check("set_gain_net")

    def get_gain_net(self, net: int) -> float:
        g = _lib.lib().leviathan_get_gain_net(self._h, int(net))
        _lib.check("get_gain_net")
        return float(g)

    def set_gain_controller_net(self, net: int, enabled: bool) -> None:
        if _lib.lib().leviathan_set_gain_controller_net(self._h, int(net), bool(enabled)) != 0:
            _lib.check("set_gain_controller_net")

    @property
    def step_count(self) -> int:
        return int(_lib.lib().leviathan_step_count(self._h))

    @property
    def num_blocks(self) -> int:
        return int(_lib.lib().leviathan_num_blocks(self._h))

    def profile_steps(self, dt: float, n: int) -> dict:
        """Times n eager steps with HIP events on the engine's stream (ms per launch)."""
        out = np.zeros(5, dtype=np.float64)
        if _lib.lib().leviathan_profile_steps(self._h, float(dt), int(n), _p(out)) != 0:
            _lib.check("profile_steps")
        return {"step_kernel_ms": float(out[0]), "finalize_kernel_ms": float(out[1]), "wall_ms": float(out[2]),
                "scatter_kernel_ms": float(out[3]), "regroup_kernel_ms": float(out[4])}

    def set_schedule(self, schedule: str, edges_per_thread: int = 8) -> None:
        if _lib.lib().leviathan_set_schedule(self._h, _lib.SCHEDULE[schedule], int(edges_per_thread)) != 0:
            _lib.check("set_schedule")

    def get_schedule(self):
        s = ctypes.c_int(0)
        e = ctypes.c_int(0)
        if _lib.lib().leviathan_get_schedule(self._h, ctypes.byref(s), ctypes.byref(e)) != 0:
            _lib.check("get_schedule")
        return ({v: k for k, v in _lib.SCHEDULE.items()}[s.value], e.value)

    @property
    def schedule_name(self) -> str:
        s, e = self.get_schedule()
        return f"{s}/{e}"

    def autotune(self, dt: float, steps: int = 10) -> dict:
        """Times every built schedule from the current state (which is restored) and keeps the fastest.
        Returns ms per step of each candidate."""
        out = np.zeros(len(_lib.PLAN_NAMES), dtype=np.float64)
        if _lib.lib().leviathan_autotune(self._h, float(dt), int(steps), _p(out)) != 0:
            _lib.check("autotune")
        return {n: float(v) for n, v in zip(_lib.PLAN_NAMES, out) if v == v}

    @property
    def scatter_steps(self) -> int:
        """Consecutive steps of step_n whose delayed phases one scatter pass gathers (two-phase schedules)."""
        k = _lib.lib().leviathan_get_scatter_steps(self._h)
        if k < 0:
            _lib.check("get_scatter_steps")
        return k

    def set_scatter_steps(self, steps: int) -> None:
        """1, 2 or 4 steps per scatter pass; results are bit-identical for every choice."""
        if _lib.lib().leviathan_set_scatter_steps(self._h, int(steps)) != 0:
            _lib.check("set_scatter_steps")

    def autotune_scatter_steps(self, dt: float, steps: int = 12) -> dict:
        """Times 1, 2 and 4 steps per scatter pass on the active schedule from the current state (restored)
        and keeps the fastest.  Returns ms per step of each (empty for the direct schedule)."""
        out = np.zeros(3, dtype=np.float64)
        if _lib.lib().leviathan_autotune_scatter_steps(self._h, float(dt), int(steps), _p(out)) != 0:
            _lib.check("autotune_scatter_steps")
        return {k: float(v) for k, v in zip((1, 2, 4), out) if v == v}

    @property
    def row_steps(self) -> int:
        """Consecutive steps of step_n one row pass covers: 2 (paired row passes) or 1."""
        k = _lib.lib().leviathan_get_row_steps(self._h)
        if k < 0:
            _lib.check("get_row_steps")
        return k

    def set_row_steps(self, steps: int) -> None:
        """1 or 2 (paired row passes, refused for a graph with delay-0 edges); results are
        bit-identical either way."""
        if _lib.lib().leviathan_set_row_steps(self._h, int(steps)) != 0:
            _lib.check("set_row_steps")

    def autotune_all(self, dt: float, steps: int = 8) -> dict:
        """The plans, then the steps per scatter pass of the fastest two-phase plan, then the plans
        again (the two-phase ones at that many steps per pass).  Each timing starts from the current
        state, which is restored.  Returns {"plans": ms per plan, "scatter_steps": ms per K,
        "plans_at_k": ms per plan at the chosen K (absent if K stayed 1)}."""
        out = {"plans": self.autotune(dt, steps)}
        two_phase = {k: v for k, v in out["plans"].items() if not k.startswith("direct")}
        if not two_phase:
            return out
        chosen = self.get_schedule()
        best = min(two_phase, key=two_phase.get)
        self.set_schedule(best.split("/")[0], int(best.split("/")[1]))
        out["scatter_steps"] = self.autotune_scatter_steps(dt, steps)
        if self.scatter_steps > 1:
            out["plans_at_k"] = self.autotune(dt, steps)
        else:
            self.set_schedule(*chosen)
        return out

    def close(self) -> None:
        if self._h is not None:
            _lib.lib().leviathan_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
