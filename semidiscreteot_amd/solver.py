"""Python mirror of the reference's SotSolver<dim, spacedim> interface
(include/SemiDiscreteOT/solvers/SotSolver.h:176-348) over the C ABI.

Method names, argument meaning and error behaviour follow the reference:
invalid measures raise (SotSolver.cc:110-112), non-convergence does not
(SotSolver.cc:207-211) and is reported by get_convergence_status().
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, fields

import numpy as np

from . import _lib as L
from .epsilon_scaling import epsilon_schedule

_THRESH = {"pointwise": 0, "integral": 1, "geometric": 2}
_CTRL = {"l1norm": 0, "componentwise": 1}
_PREC = {"fp64": 0, "fp32_lse": 1}
_LS = {"reference": 0, "noise_aware": 1}
_PC = {"none": 0, "diagonal": 1}


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


@dataclass
class SolverParameters:
    """SotParameterManager::SolverParameters (ParameterManager.h:56-72) plus
    `precision` ("fp64" | "fp32_lse"), `line_search` ("reference" |
    "noise_aware"), `preconditioner` ("none" | "diagonal": initial inverse
    Hessian diag(epsilon / nu_j) of the L-BFGS recursion) and `prune_bits`
    (value bound of the pair sums: < 0 the precision's default, 0 off),
    include/sdot.h."""
    max_iterations: int = 1000
    tolerance: float = 1.0
    solver_control_type: str = "l1norm"
    epsilon: float = 1e-3
    tau: float = 1e-8
    distance_threshold_type: str = "pointwise"
    verbose_output: bool = True
    solver_type: str = "BFGS"
    quadrature_order: int = 3
    nb_points: int = 1000
    n_threads: int = 0
    use_epsilon_scaling: bool = False
    epsilon_scaling_factor: float = 2.0
    epsilon_scaling_steps: int = 5
    use_log_sum_exp_trick: bool = False
    precision: str = "fp64"
    line_search: str = "reference"
    preconditioner: str = "none"
    prune_bits: float = -1.0

    def to_c(self) -> L.Params:
        p = L.Params()
        p.max_iterations = int(self.max_iterations)
        p.tolerance = float(self.tolerance)
        p.solver_control_type = 1 if self.solver_control_type == "componentwise" else 0
        p.epsilon = float(self.epsilon)
        p.tau = float(self.tau)
        p.distance_threshold_type = _THRESH.get(self.distance_threshold_type, 0)
        p.verbose_output = int(bool(self.verbose_output))
        p.solver_type = 0
        p.quadrature_order = int(self.quadrature_order)
        p.nb_points = int(self.nb_points)
        p.number_of_threads = int(self.n_threads)
        p.use_epsilon_scaling = int(bool(self.use_epsilon_scaling))
        p.epsilon_scaling_factor = float(self.epsilon_scaling_factor)
        p.epsilon_scaling_steps = int(self.epsilon_scaling_steps)
        p.use_log_sum_exp_trick = int(bool(self.use_log_sum_exp_trick))
        p.precision = _PREC[self.precision]
        p.line_search = _LS[self.line_search]
        p.preconditioner = _PC[self.preconditioner]
        p.prune_bits = float(self.prune_bits)
        return p

    @classmethod
    def from_c(cls, p: L.Params) -> "SolverParameters":
        inv_t = {v: k for k, v in _THRESH.items()}
        inv_p = {v: k for k, v in _PREC.items()}
        inv_l = {v: k for k, v in _LS.items()}
        inv_c = {v: k for k, v in _PC.items()}
        return cls(max_iterations=p.max_iterations, tolerance=p.tolerance,
                   solver_control_type="componentwise" if p.solver_control_type == 1 else "l1norm",
                   epsilon=p.epsilon, tau=p.tau,
                   distance_threshold_type=inv_t[p.distance_threshold_type],
                   verbose_output=bool(p.verbose_output), quadrature_order=p.quadrature_order,
                   nb_points=p.nb_points, n_threads=p.number_of_threads,
                   use_epsilon_scaling=bool(p.use_epsilon_scaling),
                   epsilon_scaling_factor=p.epsilon_scaling_factor,
                   epsilon_scaling_steps=p.epsilon_scaling_steps,
                   use_log_sum_exp_trick=bool(p.use_log_sum_exp_trick),
                   precision=inv_p[p.precision], line_search=inv_l[p.line_search],
                   preconditioner=inv_c[p.preconditioner], prune_bits=p.prune_bits)

    @classmethod
    def from_prm(cls, text: str | None = None, path: str | None = None) -> "SolverParameters":
        """Read `subsection rsot_solver` of a deal.II .prm (ParameterManager.cc:64-95)."""
        lib = L.load()
        p = L.Params()
        L.check(lib.sdot_params_default(C.byref(p)))
        if path is not None:
            L.check(lib.sdot_params_load_prm(path.encode(), C.byref(p)))
        else:
            L.check(lib.sdot_params_parse_prm((text or "").encode(), C.byref(p)))
        return cls.from_c(p)


class SotSolver:
    """SotSolver<dim> on one MI355X.  setup_source takes the flat arrays a
    deal.II caller pulls once from FEValues (quadrature points, rho*JxW, cell
    centers); see semidiscreteot_amd.mesh for generators of those arrays."""

    def __init__(self, dim: int, device: int = 0):
        self.lib = L.load()
        self.dim = int(dim)
        h = C.c_void_p()
        L.check(self.lib.sdot_ctx_create(int(device), C.byref(h)))
        self._h = h
        self.current_params = SolverParameters()
        self._n = 0
        self._nu = None
        self._last_info = None

    def close(self):
        if getattr(self, "_h", None):
            self.lib.sdot_ctx_destroy(self._h)
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

    # ---- setup (SotSolver.h:186-204) ----
    def setup_source(self, x, w, centers, nq: int):
        x, w, centers = _f64(x), _f64(w), _f64(centers)
        n_cells = centers.size // self.dim
        L.check(self.lib.sdot_set_source(self._h, self.dim, n_cells, int(nq), _p(x), _p(w),
                                         _p(centers)))
        self.n_cells, self.nq = n_cells, int(nq)

    def setup_target(self, target_points, target_density):
        y, nu = _f64(target_points), _f64(target_density)
        if y.size != nu.size * self.dim:
            raise ValueError(f"Dimension mismatch: {y.size // self.dim} points, {nu.size} weights")
        L.check(self.lib.sdot_set_target(self._h, nu.size, _p(y), _p(nu)))
        self._n = nu.size
        self._nu = nu

    def configure(self, params: SolverParameters):
        self.current_params = params
        L.check(self.lib.sdot_set_params(self._h, C.byref(params.to_c())))

    # ---- hot path ----
    def evaluate_functional(self, potential, gradient_out=None):
        """SotSolver::evaluate_functional (SotSolver.cc:217-377)."""
        psi = _f64(potential)
        g = np.zeros(self._n) if gradient_out is None else gradient_out
        F = C.c_double()
        Cg = C.c_double()
        L.check(self.lib.sdot_eval(self._h, _p(psi), _p(g), C.byref(F), C.byref(Cg)))
        return F.value if gradient_out is not None else (F.value, g)

    def evaluate_functional_R(self, R, potential):
        psi = _f64(potential)
        g = np.zeros(self._n)
        F = C.c_double()
        Cg = C.c_double()
        L.check(self.lib.sdot_eval_R(self._h, float(R), _p(psi), _p(g), C.byref(F), C.byref(Cg)))
        return F.value, g, Cg.value

    def evaluate_weighted_barycenters(self, potential, params: SolverParameters | None = None):
        """SotSolver::evaluate_weighted_barycenters with the Euclidean distance
        (SotSolver.cc:585-610, compute_weighted_barycenters_euclidean :871-971):
        per target the first moment of its column of the transport plan over
        nu_j, with the threshold from the stored state.  `params` is accepted
        for the reference's signature; the evaluation uses the configured
        epsilon, as the reference uses the last solve's.  Returns [N, dim]."""
        if self._n == 0:
            raise RuntimeError("Invalid measures configuration")
        out = np.zeros(self._n * self.dim)
        L.check(self.lib.sdot_barycenters(self._h, _p(_f64(potential)), _p(out)))
        return out.reshape(self._n, self.dim)

    def compute_weighted_barycenters_R(self, R, potential):
        """The barycenters at an explicit threshold R (no state involved)."""
        out = np.zeros(self._n * self.dim)
        L.check(self.lib.sdot_barycenters_R(self._h, float(R), _p(_f64(potential)), _p(out)))
        return out.reshape(self._n, self.dim)

    def get_potential_conditioned_density(self, points, density, potential, potential_indices,
                                          R=None):
        """SotSolver::get_potential_conditioned_density (SotSolver.cc:973-1046)
        at `points` (the reference's DoF support points, [n, dim]) with the
        source density there: row k is rho(p) exp(e_pi - M_p) / S_p for
        i = potential_indices[k], 0 where i is not within the threshold of p.
        The threshold is the last distance threshold (the reference's
        current_distance_threshold) unless R is given; epsilon is the configured
        one.  Returns [len(potential_indices), n]."""
        if self._n == 0:
            raise RuntimeError("Invalid measures configuration")
        pts = _f64(points).reshape(-1, self.dim)
        n = len(pts)
        rho = _f64(density).ravel()
        if rho.size != n:
            raise ValueError("density must have one value per point")
        idx = np.ascontiguousarray(np.asarray(potential_indices, dtype=np.int64).ravel())
        out = np.zeros((idx.size, n))
        thr = self.get_last_distance_threshold() if R is None else float(R)
        L.check(self.lib.sdot_conditional_density(
            self._h, _p(_f64(potential)), thr, n, _p(pts), _p(rho), idx.size,
            idx.ctypes.data_as(L._i64p), _p(out)))
        return out

    def compute_hessian(self, potential):
        psi = _f64(potential)
        H = np.zeros((self._n, self._n))
        L.check(self.lib.sdot_hessian(self._h, _p(psi), _p(H), self._n))
        return H

    def solve(self, potential=None, params: SolverParameters | None = None):
        """SotSolver::solve (SotSolver.cc:105-215).  Returns the potential."""
        if params is not None:
            self.configure(params)
        if self._n == 0:
            raise RuntimeError("Invalid measures configuration")
        psi = np.zeros(self._n) if potential is None or np.size(potential) != self._n \
            else _f64(potential).copy()
        info = L.SolveInfo()
        L.check(self.lib.sdot_solve(self._h, _p(psi), C.byref(info)))
        self._last_info = {"status": L.SOLVE_STATUS[info.status], "iterations": info.iterations,
                           "evaluations": info.evaluations, "functional": info.functional,
                           "check_value": info.check_value, "seconds": info.seconds}
        if potential is not None and np.size(potential) == self._n and \
                isinstance(potential, np.ndarray) and potential.dtype == np.float64:
            potential[:] = psi
        return psi

    def solve_epsilon_scaling(self, potential=None, params: SolverParameters | None = None):
        """The epsilon-scaling loop SemiDiscreteOT::run_sot wraps around
        SotSolver::solve (SemiDiscreteOT.cc:877-918): solves at
        epsilon * factor^(steps-1-i), i = 0..steps-1 (epsilon_scaling.epsilon_schedule;
        factor <= 1 -> 2 and steps 0 -> 1 as EpsilonScalingHandler.cc:18-27), each to the tolerance, carrying the potential
        over; a stage that does not converge does not stop the loop.  Without
        use_epsilon_scaling it is a single solve.  Returns (psi, [stage info])."""
        if params is not None:
            self.configure(params)
        p = self.current_params
        if p.use_epsilon_scaling:
            seq = epsilon_schedule(p.epsilon, p.epsilon_scaling_factor,
                                   p.epsilon_scaling_steps) or [p.epsilon]
        else:
            seq = [p.epsilon]
        psi = None if potential is None else _f64(potential).copy()
        infos = []
        try:
            for eps in seq:
                stage = SolverParameters(**{f.name: getattr(p, f.name) for f in fields(p)})
                stage.epsilon = eps
                psi = self.solve(psi, stage)
                infos.append(dict(self.get_last_solve_info(), epsilon=stage.epsilon))
        finally:
            self.configure(p)
        return psi, infos

    # ---- getters (SotSolver.h:240-262) ----
    def _state(self):
        st = L.State()
        L.check(self.lib.sdot_get_state(self._h, C.byref(st)))
        return st

    def get_last_functional_value(self):
        return self._state().last_functional

    def get_last_iteration_count(self):
        return self._state().last_iteration_count

    def get_convergence_status(self):
        return bool(self._state().converged)

    def get_last_distance_threshold(self):
        return self._state().last_distance_threshold

    def get_C_global(self):
        return self._state().C_global

    def set_distance_threshold(self, threshold):
        L.check(self.lib.sdot_set_distance_threshold(self._h, float(threshold)))

    def get_last_solve_info(self):
        return self._last_info

    def compute_covering_radius(self):
        r = C.c_double()
        L.check(self.lib.sdot_covering_radius(self._h, C.byref(r)))
        return r.value

    def distance_threshold(self, potential):
        R = C.c_double()
        L.check(self.lib.sdot_distance_threshold(self._h, _p(_f64(potential)), C.byref(R)))
        return R.value

    def candidate_stats(self, R):
        cnt = np.zeros(self.n_cells, dtype=np.int64)
        h = np.zeros(self.n_cells, dtype=np.uint64)
        pairs = C.c_int64()
        L.check(self.lib.sdot_candidate_stats(self._h, float(R),
                                              cnt.ctypes.data_as(C.POINTER(C.c_int64)),
                                              h.ctypes.data_as(C.POINTER(C.c_uint64)),
                                              C.byref(pairs)))
        return cnt, h, pairs.value

    def q_log_partition(self, R, potential):
        out = np.zeros(self.n_cells * self.nq)
        L.check(self.lib.sdot_q_log_partition(self._h, float(R), _p(_f64(potential)), _p(out)))
        return out

    def discarded_mass(self, R, potential, w, logZ_full=None):
        """Mass the truncation radius R drops: sum_q w_q (1 - Z_q(R)/Z_q(inf))."""
        if logZ_full is None:
            logZ_full = self.q_log_partition(np.finfo(float).max, potential)
        lz = self.q_log_partition(R, potential)
        frac = np.where(np.isfinite(lz), -np.expm1(np.minimum(lz - logZ_full, 0.0)), 1.0)
        return float(np.sum(np.asarray(w) * frac))

    # ---- device-resident evaluation and profiling ----
    def eval_device(self, R, d_psi: int, d_grad: int, d_FC: int):
        L.check(self.lib.sdot_eval_device(self._h, float(R), C.c_void_p(d_psi),
                                          C.c_void_p(d_grad), C.c_void_p(d_FC)))

    def eval_device_noalloc(self, R, d_psi: int, d_grad: int, d_FC: int):
        """eval_device with every per-evaluation buffer taken from the
        library's fixed 128 MiB device array (no device allocation)."""
        L.check(self.lib.sdot_eval_device_noalloc(self._h, float(R), C.c_void_p(d_psi),
                                                  C.c_void_p(d_grad), C.c_void_p(d_FC)))

    def synchronize(self):
        L.check(self.lib.sdot_synchronize(self._h))

    def profile_enable(self, on=True):
        L.check(self.lib.sdot_profile_enable(self._h, int(bool(on))))

    def profile_get(self):
        p = L.Profile()
        L.check(self.lib.sdot_profile_get(self._h, C.byref(p)))
        return {f[0]: getattr(p, f[0]) for f in L.Profile._fields_}
