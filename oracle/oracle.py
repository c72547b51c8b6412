"""ctypes binding of the FP64 CPU oracle (liboracle_sdot.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the checker or the timed CPU
baseline.  The product package (semidiscreteot_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_sdot.so")

THRESH = {"pointwise": 0, "integral": 1, "geometric": 2}
CTRL = {"l1norm": 0, "componentwise": 1}
STATUS = {0: "converged", 1: "max_iterations", 2: "linesearch_fail", 3: "nonfinite"}

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


class _Source(C.Structure):
    _fields_ = [("dim", C.c_int), ("n_cells", C.c_int64), ("nq", C.c_int),
                ("x", _dp), ("rho", _dp), ("jxw", _dp), ("center", _dp)]


class _Target(C.Structure):
    _fields_ = [("n", C.c_int64), ("y", _dp), ("nu", _dp)]


class _Params(C.Structure):
    _fields_ = [("epsilon", C.c_double), ("tau", C.c_double), ("tolerance", C.c_double),
                ("use_log_sum_exp", C.c_int), ("threshold_type", C.c_int),
                ("control_type", C.c_int), ("max_iterations", C.c_int),
                ("n_threads", C.c_int), ("faithful_copier", C.c_int)]


class _State(C.Structure):
    _fields_ = [("covering_radius", C.c_double), ("min_target_density", C.c_double),
                ("C_global", C.c_double), ("global_functional", C.c_double),
                ("last_threshold", C.c_double)]


class _Info(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_int), ("evaluations", C.c_int),
                ("functional", C.c_double), ("check_value", C.c_double)]


_FUN = C.CFUNCTYPE(C.c_double, C.c_void_p, _dp, _dp)

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_gauss_legendre01.argtypes = [C.c_int, _dp, _dp]
        L.orc_structured_source.argtypes = [C.c_int, C.POINTER(C.c_int), _dp, _dp, C.c_int,
                                            _dp, _dp, _dp, _dp, _dp]
        L.orc_simplex_source.argtypes = [C.c_int, C.c_int64, _dp, C.c_int64, _i64p, C.c_int,
                                         _dp, _dp, _dp, _dp, _dp]
        L.orc_cells_source.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int64, _dp, C.c_int64,
                                       _i64p, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.orc_normalize_density.argtypes = [C.c_int64, _dp, _dp]
        L.orc_normalize_density.restype = C.c_double
        L.orc_covering_radius.argtypes = [C.POINTER(_Source), C.POINTER(_Target)]
        L.orc_covering_radius.restype = C.c_double
        L.orc_distance_threshold.argtypes = [C.POINTER(_Params), C.POINTER(_State),
                                             C.POINTER(_Source), C.POINTER(_Target), _dp]
        L.orc_distance_threshold.restype = C.c_double
        L.orc_evaluate.argtypes = [C.POINTER(_Source), C.POINTER(_Target), C.POINTER(_Params),
                                   C.POINTER(_State), _dp, _dp, _dp]
        L.orc_evaluate_R.argtypes = [C.POINTER(_Source), C.POINTER(_Target), C.POINTER(_Params),
                                     C.c_double, _dp, _dp, _dp, _dp]
        L.orc_candidate_stats.argtypes = [C.POINTER(_Source), C.POINTER(_Target), C.c_double,
                                          _i64p, C.POINTER(C.c_uint64)]
        L.orc_discarded_fraction.argtypes = [C.POINTER(_Source), C.POINTER(_Target),
                                             C.POINTER(_Params), C.c_double, _dp, C.c_int64,
                                             _i64p, _dp]
        L.orc_hessian.argtypes = [C.POINTER(_Source), C.POINTER(_Target), C.POINTER(_Params),
                                  C.c_double, _dp, _dp]
        L.orc_solve.argtypes = [C.POINTER(_Source), C.POINTER(_Target), C.POINTER(_Params),
                                C.POINTER(_State), _dp, C.POINTER(_Info)]
        L.orc_bfgs_generic.argtypes = [_FUN, C.c_void_p, C.c_int64, _dp, C.c_int, C.c_double,
                                       C.POINTER(_Info)]
        L.orc_barycenters_R.argtypes = [C.POINTER(_Source), C.POINTER(_Target), C.POINTER(_Params),
                                        C.c_double, _dp, _dp]
        L.orc_conditional_density.argtypes = [C.POINTER(_Target), C.c_int, _dp, C.c_double,
                                              C.c_double, C.c_int64, _dp, _dp, C.c_int64, _i64p,
                                              _dp]
        L.orc_transport_map.argtypes = [C.c_int, C.POINTER(_Target), C.c_int, _dp, C.c_double,
                                        C.c_double, C.c_int64, _dp, _dp]
        L.orc_softmax_refinement.argtypes = [C.POINTER(_Source), C.c_int64, _dp,
                                             C.POINTER(_Target), _dp, C.c_int64, _i64p, _i64p,
                                             C.c_double, C.c_double, C.c_int, C.c_int, _dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def gauss_legendre01(n):
    p = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss_legendre01(n, _p(p), _p(w))
    return p, w


class Source:
    """Flat source arrays as FEValues would hand them to local_assemble."""

    def __init__(self, dim, nq, x, rho, jxw, center):
        self.dim, self.nq = int(dim), int(nq)
        self.x, self.rho, self.jxw, self.center = _f64(x), _f64(rho), _f64(jxw), _f64(center)
        self.n_cells = self.center.size // self.dim
        assert self.x.size == self.n_cells * self.nq * self.dim
        self._s = _Source(self.dim, self.n_cells, self.nq, _p(self.x), _p(self.rho),
                          _p(self.jxw), _p(self.center))

    @property
    def w(self):
        return self.rho * self.jxw


class Target:
    def __init__(self, y, nu):
        self.y, self.nu = _f64(y), _f64(nu)
        self.n = self.nu.size
        self._t = _Target(self.n, _p(self.y), _p(self.nu))


def structured_source(dim, n_sub, lo, hi, qorder, nodal_rho=None, normalize=True):
    n_sub = list(n_sub)
    ns = n_sub + [1] * (3 - len(n_sub))
    n_cells = int(np.prod(ns[:dim]))
    nq = qorder ** dim
    nverts = int(np.prod([n + 1 for n in ns[:dim]]))
    rho_n = None if nodal_rho is None else _f64(nodal_rho).copy()
    if rho_n is not None:
        assert rho_n.size == nverts

    def run(rn):
        x = np.zeros(n_cells * nq * dim)
        rho = np.zeros(n_cells * nq)
        jxw = np.zeros(n_cells * nq)
        c = np.zeros(n_cells * dim)
        rc = lib().orc_structured_source(dim, (C.c_int * 3)(*ns), _p(_f64(lo)), _p(_f64(hi)),
                                         qorder, None if rn is None else _p(rn),
                                         _p(x), _p(rho), _p(jxw), _p(c))
        assert rc == 0
        return x, rho, jxw, c

    x, rho, jxw, c = run(rho_n)
    if normalize:
        L = lib().orc_normalize_density(rho.size, _p(rho), _p(jxw))
        if rho_n is None:
            rho_n = np.ones(nverts)
        rho_n = rho_n / L
        x, rho, jxw, c = run(rho_n)
    return Source(dim, nq, x, rho, jxw, c)


def simplex_source(dim, vertices, cells, qorder, nodal_rho=None, normalize=True):
    vertices = _f64(vertices)
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    n_cells = cells.size // (dim + 1)
    nq = {1: 1, 2: dim + 1}[qorder]
    rho_n = np.ones(vertices.size // dim) if nodal_rho is None else _f64(nodal_rho).copy()

    def run(rn):
        x = np.zeros(n_cells * nq * dim)
        rho = np.zeros(n_cells * nq)
        jxw = np.zeros(n_cells * nq)
        c = np.zeros(n_cells * dim)
        rc = lib().orc_simplex_source(dim, vertices.size // dim, _p(vertices), n_cells,
                                      cells.ctypes.data_as(_i64p), qorder, _p(rn),
                                      _p(x), _p(rho), _p(jxw), _p(c))
        assert rc == nq
        return x, rho, jxw, c

    x, rho, jxw, c = run(rho_n)
    if normalize:
        L = lib().orc_normalize_density(rho.size, _p(rho), _p(jxw))
        rho_n = rho_n / L
        x, rho, jxw, c = run(rho_n)
    return Source(dim, nq, x, rho, jxw, c)


def cells_source(q1, cdim, sdim, vertices, cells, qorder, nodal_rho=None, normalize=True):
    """General Q1 (q1=True) or P1 cells of dimension cdim in dimension sdim;
    the Source has dim = sdim."""
    vertices = _f64(vertices)
    cells = np.ascontiguousarray(cells, dtype=np.int64)
    nv = 2 ** cdim if q1 else cdim + 1
    n_cells = cells.size // nv
    nq = qorder ** cdim if q1 else {1: 1, 2: cdim + 1}[qorder]
    rho_n = np.ones(vertices.size // sdim) if nodal_rho is None else _f64(nodal_rho).copy()

    def run(rn):
        x = np.zeros(n_cells * nq * sdim)
        rho = np.zeros(n_cells * nq)
        jxw = np.zeros(n_cells * nq)
        c = np.zeros(n_cells * sdim)
        rc = lib().orc_cells_source(int(bool(q1)), cdim, sdim, vertices.size // sdim,
                                    _p(vertices), n_cells, cells.ctypes.data_as(_i64p), qorder,
                                    _p(rn), _p(x), _p(rho), _p(jxw), _p(c))
        assert rc == nq
        return x, rho, jxw, c

    x, rho, jxw, c = run(rho_n)
    if normalize:
        L = lib().orc_normalize_density(rho.size, _p(rho), _p(jxw))
        rho_n = rho_n / L
        x, rho, jxw, c = run(rho_n)
    return Source(sdim, nq, x, rho, jxw, c)


def params(epsilon=1e-3, tau=1e-8, tolerance=1.0, use_lse=False, threshold="pointwise",
           control="l1norm", max_iterations=1000, n_threads=0, faithful=True):
    return _Params(float(epsilon), float(tau), float(tolerance), int(bool(use_lse)),
                   THRESH[threshold], CTRL[control], int(max_iterations), int(n_threads),
                   int(bool(faithful)))


def new_state():
    return _State(0.0, 0.0, 0.0, 0.0, 0.0)


def evaluate_R(src, tgt, prm, R, psi):
    psi = _f64(psi)
    g = np.zeros(tgt.n)
    F = C.c_double()
    Cg = C.c_double()
    rc = lib().orc_evaluate_R(C.byref(src._s), C.byref(tgt._t), C.byref(prm), float(R), _p(psi),
                              _p(g), C.byref(F), C.byref(Cg))
    assert rc == 0
    return F.value, g, Cg.value


def evaluate(src, tgt, prm, state, psi):
    psi = _f64(psi)
    g = np.zeros(tgt.n)
    F = C.c_double()
    rc = lib().orc_evaluate(C.byref(src._s), C.byref(tgt._t), C.byref(prm), C.byref(state),
                            _p(psi), _p(g), C.byref(F))
    assert rc == 0
    return F.value, g


def distance_threshold(src, tgt, prm, state, psi):
    return lib().orc_distance_threshold(C.byref(prm), C.byref(state), C.byref(src._s),
                                        C.byref(tgt._t), _p(_f64(psi)))


def covering_radius(src, tgt):
    return lib().orc_covering_radius(C.byref(src._s), C.byref(tgt._t))


def candidate_stats(src, tgt, R):
    cnt = np.zeros(src.n_cells, dtype=np.int64)
    h = np.zeros(src.n_cells, dtype=np.uint64)
    lib().orc_candidate_stats(C.byref(src._s), C.byref(tgt._t), float(R),
                              cnt.ctypes.data_as(_i64p), h.ctypes.data_as(C.POINTER(C.c_uint64)))
    return cnt, h


def discarded_fraction(src, tgt, prm, R, psi, q_list):
    q = np.ascontiguousarray(q_list, dtype=np.int64)
    out = np.zeros(q.size)
    lib().orc_discarded_fraction(C.byref(src._s), C.byref(tgt._t), C.byref(prm), float(R),
                                 _p(_f64(psi)), q.size, q.ctypes.data_as(_i64p), _p(out))
    return out


def hessian(src, tgt, prm, R, psi):
    H = np.zeros((tgt.n, tgt.n))
    rc = lib().orc_hessian(C.byref(src._s), C.byref(tgt._t), C.byref(prm), float(R),
                           _p(_f64(psi)), _p(H))
    assert rc == 0
    return H


def solve(src, tgt, prm, psi0=None, state=None):
    psi = np.zeros(tgt.n) if psi0 is None else _f64(psi0).copy()
    st = new_state() if state is None else state
    info = _Info()
    lib().orc_solve(C.byref(src._s), C.byref(tgt._t), C.byref(prm), C.byref(st), _p(psi),
                    C.byref(info))
    return psi, {"status": STATUS.get(info.status, info.status), "iterations": info.iterations,
                 "evaluations": info.evaluations, "functional": info.functional,
                 "check_value": info.check_value, "state": st}


def bfgs_generic(fun, x0, max_iterations=1000, tol=1e-10):
    """Minimise fun(x)->(f, g) with the oracle's L-BFGS (textbook pinning)."""
    x = _f64(x0).copy()
    n = x.size

    def cb(_ctx, xp, gp):
        xv = np.ctypeslib.as_array(xp, shape=(n,)).copy()
        f, g = fun(xv)
        np.ctypeslib.as_array(gp, shape=(n,))[:] = g
        return float(f)

    cfun = _FUN(cb)
    info = _Info()
    lib().orc_bfgs_generic(cfun, None, n, _p(x), int(max_iterations), float(tol), C.byref(info))
    return x, {"status": STATUS.get(info.status, info.status), "iterations": info.iterations,
               "evaluations": info.evaluations, "functional": info.functional}


def children_csr(children):
    """List of child-index lists (one per coarse point) -> (offsets, flat)."""
    off = np.zeros(len(children) + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(c) for c in children])
    flat = np.ascontiguousarray(np.concatenate([np.asarray(c, dtype=np.int64) for c in children])
                                if len(children) else np.zeros(0, dtype=np.int64), dtype=np.int64)
    return off, flat


def softmax_refinement(src, y_fine, coarse, psi_coarse, children, lam, threshold, use_lse=True,
                       n_threads=0):
    """SoftmaxRefinement::compute_refinement (SoftmaxRefinement.cc:200-280); `children`
    is the level's list of child-index lists."""
    y_fine = _f64(y_fine)
    n_fine = y_fine.size // src.dim
    off, flat = children_csr(children)
    out = np.zeros(n_fine)
    rc = lib().orc_softmax_refinement(C.byref(src._s), n_fine, _p(y_fine), C.byref(coarse._t),
                                      _p(_f64(psi_coarse)), len(children),
                                      off.ctypes.data_as(_i64p),
                                      (flat if flat.size else np.zeros(1, dtype=np.int64)
                                       ).ctypes.data_as(_i64p),
                                      float(lam), float(threshold), int(bool(use_lse)),
                                      int(n_threads), _p(out))
    assert rc == 0
    return out


def barycenters_R(src, tgt, prm, R, psi):
    """compute_weighted_barycenters_euclidean (SotSolver.cc:871-971) at threshold R."""
    out = np.zeros(tgt.n * src.dim)
    rc = lib().orc_barycenters_R(C.byref(src._s), C.byref(tgt._t), C.byref(prm), float(R),
                                 _p(_f64(psi)), _p(out))
    assert rc == 0
    return out.reshape(tgt.n, src.dim)


def conditional_density(tgt, dim, psi, eps, R, points, density, indices):
    """get_potential_conditioned_density (SotSolver.cc:973-1046); [n_idx, n_points]."""
    pts = _f64(points).ravel()
    n = pts.size // dim
    idx = np.ascontiguousarray(indices, dtype=np.int64)
    out = np.zeros(idx.size * n)
    lib().orc_conditional_density(C.byref(tgt._t), dim, _p(_f64(psi)), float(eps), float(R), n,
                                  _p(pts), _p(_f64(density)), idx.size, idx.ctypes.data_as(_i64p),
                                  _p(out))
    return out.reshape(idx.size, n)


def transport_map(strategy, tgt, dim, psi, eps, radius, points):
    """Modal (0) / barycentric (1) compute_map (OptimalTransportPlan.cc:99-310)."""
    pts = _f64(points).ravel()
    n = pts.size // dim
    out = np.zeros(n * dim)
    lib().orc_transport_map(int(strategy), C.byref(tgt._t), dim, _p(_f64(psi)), float(eps),
                            float(radius), n, _p(pts), _p(out))
    return out.reshape(n, dim)
