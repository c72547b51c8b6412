"""Batched solve queue over the C ABI (include/sdot.h, "batched serving").

Many SotSolver::solve-style solves (SotSolver.cc:105-215: L-BFGS m = 5, the
deal.II line search, the l1norm / componentwise stopping tests) share one
resident source mesh and advance together, one evaluation of each active
solve per step.  submit() and poll() may be called from any thread; step()
from one thread at a time.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .solver import _PC, SotSolver, _f64, _p

_PREC = {"fp64": 0, "fp32_lse": 1}
_CTRL = {"l1norm": 0, "componentwise": 1}
_LS = {"reference": 0, "noise_aware": 1}


@dataclass
class ServeResult:
    id: int
    status: str
    iterations: int
    evaluations: int
    functional: float
    check_value: float
    rel_marginal_error: float
    t_submit: float
    t_admit: float
    t_done: float

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit


class SotServer:
    """Serving queue bound to the source measure of `solver` (a SotSolver
    whose setup_source has run).  The solver's source stays fixed while the
    server exists."""

    def __init__(self, solver: SotSolver, precision: str = "fp32_lse", max_slots: int = 256,
                 hbm_budget_bytes: int = 16 << 30, use_graph: bool = True,
                 pass_targets: int = 0, grad_qpoints: int = 0, polish_factor: float = 2.0):
        self.lib = L.load()
        self.dim = solver.dim
        self._solver = solver  # keeps the source context alive
        cfg = L.ServerConfig()
        L.check(self.lib.sdot_server_config_default(C.byref(cfg)))
        cfg.precision = _PREC[precision]
        cfg.max_slots = int(max_slots)
        cfg.hbm_budget_bytes = int(hbm_budget_bytes)
        cfg.use_graph = int(bool(use_graph))
        cfg.pass_targets = int(pass_targets)
        cfg.grad_qpoints = int(grad_qpoints)
        cfg.polish_factor = float(polish_factor)
        h = C.c_void_p()
        L.check(self.lib.sdot_server_create(solver._h, C.byref(cfg), C.byref(h)))
        self._h = h
        self._n = {}

    def close(self):
        if getattr(self, "_h", None):
            self.lib.sdot_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def submit(self, y, nu, epsilon: float, tolerance: float = 1e-6,
               control: str = "componentwise", max_iterations: int = 1000,
               radius: float | None = None, psi0=None,
               line_search: str = "noise_aware", epsilon_scaling_steps: int = 0,
               epsilon_scaling_factor: float = 2.0, preconditioner: str = "none") -> int:
        y, nu = _f64(y), _f64(nu)
        if y.size != nu.size * self.dim:
            raise ValueError(f"Dimension mismatch: {y.size // self.dim} points, {nu.size} weights")
        q = L.Request()
        q.n = nu.size
        q.y = _p(y)
        q.nu = _p(nu)
        if psi0 is not None:
            psi0 = _f64(psi0)
            if psi0.size != nu.size:
                raise ValueError("psi0 has the wrong size")
            q.psi0 = _p(psi0)
        q.epsilon = float(epsilon)
        q.tolerance = float(tolerance)
        q.solver_control_type = _CTRL[control]
        q.max_iterations = int(max_iterations)
        q.radius = 0.0 if radius is None else float(radius)
        q.line_search = _LS[line_search]
        q.epsilon_scaling_steps = int(epsilon_scaling_steps)
        q.epsilon_scaling_factor = float(epsilon_scaling_factor)
        q.preconditioner = _PC[preconditioner]
        rid = C.c_int64(-1)
        L.check(self.lib.sdot_server_submit(self._h, C.byref(q), C.byref(rid)))
        self._n[rid.value] = nu.size
        return rid.value

    def step(self):
        """One batched step; returns (active solves, queued requests)."""
        na, nq = C.c_int32(), C.c_int32()
        L.check(self.lib.sdot_server_step(self._h, C.byref(na), C.byref(nq)))
        return na.value, nq.value

    def poll(self, rid: int):
        """(ServeResult, psi) once request `rid` has finished, else None."""
        done = C.c_int32()
        res = L.Result()
        psi = np.zeros(self._n[rid])
        L.check(self.lib.sdot_server_poll(self._h, int(rid), C.byref(done), C.byref(res), _p(psi)))
        if not done.value:
            return None
        del self._n[rid]
        return ServeResult(res.id, L.SOLVE_STATUS[res.status], res.iterations, res.evaluations,
                           res.functional, res.check_value, res.rel_marginal_error,
                           res.t_submit, res.t_admit, res.t_done), psi

    def clock(self) -> float:
        t = C.c_double()
        L.check(self.lib.sdot_server_clock(self._h, C.byref(t)))
        return t.value

    def request_bytes(self, n: int, radius: float | None = None) -> int:
        b = C.c_int64()
        L.check(self.lib.sdot_server_request_bytes(self._h, int(n),
                                                   0.0 if radius is None else float(radius),
                                                   C.byref(b)))
        return b.value

    def autotune(self):
        """Pick (pass_targets, grad_qpoints) by timing; call before submitting."""
        a, b = C.c_int32(), C.c_int32()
        L.check(self.lib.sdot_server_autotune(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def run(self, ids=None):
        """Step until every listed (default: every submitted) request is done;
        returns {id: (ServeResult, psi)}."""
        want = set(self._n) if ids is None else set(ids)
        out = {}
        while want:
            self.step()
            for rid in list(want):
                r = self.poll(rid)
                if r is not None:
                    out[rid] = r
                    want.discard(rid)
        return out
