"""Multilevel callers of the solver: the softmax potential transfer between
target levels (SoftmaxRefinement, src/solvers/SoftmaxRefinement.cc:61-280,
on the GPU through sdot_softmax_refinement), the coarse-to-fine assignment
(SemiDiscreteOT::assign_potentials_by_hierarchy, SemiDiscreteOT.cc:748-830),
the target point-cloud hierarchy (PointCloudHierarchyManager::generateHierarchy,
src/core/PointCloudHierarchy.cc:126-300) and the target-only and combined
source/target multilevel drivers (SemiDiscreteOT::run_target_multilevel,
SemiDiscreteOT.cc:966-1323, run_combined_multilevel, :1559-1844).

Each level's solve is SotSolver::solve on the GPU; the drivers only move
potentials and parameters between levels, as the reference's do.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import sys
import time
from dataclasses import dataclass, field, fields

import numpy as np

from . import _lib as L
from . import formats
from .epsilon_scaling import level_schedules, stages_of_level
from .solver import SolverParameters, SotSolver, _f64, _p


def children_csr(children) -> tuple[np.ndarray, np.ndarray]:
    """List of child-index lists (one per coarse point) -> (offsets, flat)."""
    off = np.zeros(len(children) + 1, dtype=np.int64)
    if len(children):
        off[1:] = np.cumsum([len(c) for c in children])
    flat = np.zeros(int(off[-1]), dtype=np.int64)
    for i, c in enumerate(children):
        flat[off[i]: off[i + 1]] = c
    return off, flat


class SoftmaxRefinement:
    """SoftmaxRefinement<dim> (SoftmaxRefinement.h; constructor
    SoftmaxRefinement.cc:4-27) bound to the source measure of `solver` (the
    reference passes the source DoFHandler, mapping, FE, density and
    quadrature order), with the distance threshold and the LSE flag."""

    def __init__(self, solver: SotSolver, distance_threshold: float,
                 use_log_sum_exp_trick: bool = True):
        self.solver = solver
        self.current_distance_threshold = float(distance_threshold)
        self.use_log_sum_exp_trick = bool(use_log_sum_exp_trick)

    def compute_refinement(self, target_points_fine, target_density_fine, target_points_coarse,
                           target_density_coarse, potential_coarse, regularization_param: float,
                           level: int, child_indices) -> np.ndarray:
        """SoftmaxRefinement::compute_refinement (SoftmaxRefinement.cc:200-280).
        child_indices[level][c] lists the fine children of coarse point c.  An
        invalid level leaves every fine potential 0 after an error message
        (SoftmaxRefinement.cc:103-107).  target_density_fine is not read by the
        algorithm."""
        dim = self.solver.dim
        yf = _f64(target_points_fine).ravel()
        yc = _f64(target_points_coarse).ravel()
        nuc = _f64(target_density_coarse).ravel()
        psic = _f64(potential_coarse).ravel()
        n_fine = yf.size // dim
        if yc.size != nuc.size * dim or psic.size != nuc.size:
            raise ValueError("coarse points, densities and potential differ in size")
        out = np.zeros(n_fine)
        if level < 0 or level >= len(child_indices):
            print(f"softmax transfer: no child lists for level {level} "
                  f"({len(child_indices)} levels known); fine potentials left at 0", file=sys.stderr)
            return out
        off, flat = children_csr(child_indices[level])
        lib = L.load()
        L.check(lib.sdot_softmax_refinement(
            self.solver._h, n_fine, _p(yf) if n_fine else None, nuc.size, _p(yc), _p(nuc),
            _p(psic), len(off) - 1, off.ctypes.data_as(C.POINTER(C.c_int64)),
            flat.ctypes.data_as(C.POINTER(C.c_int64)) if flat.size else None,
            float(regularization_param), self.current_distance_threshold,
            int(self.use_log_sum_exp_trick), _p(out) if n_fine else None))
        return out


def assign_potentials_by_hierarchy(solver: SotSolver, coarse_level: int, fine_level: int,
                                   prev_potentials, fine_points, coarse_points, coarse_density,
                                   child_indices, use_softmax: bool, distance_threshold: float,
                                   epsilon: float, use_log_sum_exp_trick: bool) -> np.ndarray:
    """SemiDiscreteOT::assign_potentials_by_hierarchy (SemiDiscreteOT.cc:748-830):
    the same level copies; otherwise the softmax refinement or, without it,
    every child takes its parent's potential (fine points without a parent
    keep 0)."""
    prev = _f64(prev_potentials).ravel()
    if coarse_level < 0 or fine_level < 0:
        print(f"potential transfer: negative level ({coarse_level} -> {fine_level})",
              file=sys.stderr)
        return None
    if coarse_level == fine_level:
        return prev.copy()
    n_fine = _f64(fine_points).size // solver.dim
    if use_softmax:
        ref = SoftmaxRefinement(solver, distance_threshold, use_log_sum_exp_trick)
        return ref.compute_refinement(fine_points, None, coarse_points, coarse_density, prev,
                                      epsilon, fine_level, child_indices)
    out = np.zeros(n_fine)
    for j, children in enumerate(child_indices[fine_level][: prev.size]):
        for ch in children:
            out[ch] = prev[j]
    return out


# ---------------------------------------------------------------------------
# target point-cloud hierarchy
# ---------------------------------------------------------------------------

def kmeans(points: np.ndarray, k: int, rng: np.random.Generator, max_iter: int = 100):
    """Lloyd's k-means from a k-means++ start (the algorithm of the dkm
    library the reference calls, PointCloudHierarchy.cc:96-117), seeded.
    Returns (centers [k, dim], assignment [n])."""
    from scipy.spatial import cKDTree

    n, dim = points.shape
    cols = [np.ascontiguousarray(points[:, d]) for d in range(dim)]
    tmp = np.empty(n)

    def dist2(c, out):
        np.subtract(cols[0], c[0], out=out)
        np.square(out, out=out)
        for d in range(1, dim):
            np.subtract(cols[d], c[d], out=tmp)
            np.square(tmp, out=tmp)
            out += tmp
        return out

    centers = np.empty((k, dim))
    centers[0] = points[rng.integers(n)]
    d2 = dist2(centers[0], np.empty(n))
    cand = np.empty(n)
    for i in range(1, k):
        # D^2 sampling: the next center with probability proportional to the
        # squared distance to the nearest chosen center
        cum = np.cumsum(d2)
        j = int(np.searchsorted(cum, rng.random() * cum[-1], side="right")) if cum[-1] > 0 \
            else int(rng.integers(n))
        centers[i] = points[min(j, n - 1)]
        np.minimum(d2, dist2(centers[i], cand), out=d2)
    assign = None
    for _ in range(max_iter):
        _, a = cKDTree(centers).query(points)
        if assign is not None and np.array_equal(a, assign):
            break
        assign = a
        cnt = np.bincount(assign, minlength=k)
        for d in range(points.shape[1]):
            s = np.bincount(assign, weights=points[:, d], minlength=k)
            centers[:, d] = np.where(cnt > 0, s / np.maximum(cnt, 1), centers[:, d])
    return centers, assign


def hierarchy_level_sizes(n_points: int, min_points: int, max_points: int) -> list[int]:
    """Level sizes of PointCloudHierarchyManager (PointCloudHierarchy.cc:60-74,
    :150-156): level 0 all points, level 1 min(n, max_points), then a factor 4
    per level down to min_points, at most 5 levels."""
    level1 = min(n_points, max_points)
    n_levels = min(5, int(math.log(level1 // min_points) / math.log(4.0)) + 2) \
        if level1 // min_points > 0 else 1
    n_levels = max(1, n_levels)
    sizes = []
    for lev in range(n_levels):
        if lev == 0:
            sizes.append(n_points)
        elif lev == 1:
            sizes.append(level1)
        else:
            sizes.append(max(int(level1 / 4.0 ** (lev - 1)), min_points))
    return sizes


def generate_target_hierarchy(points, densities=None, output_dir: str | None = None,
                              min_points: int = 100, max_points: int = 1000,
                              sizes: list[int] | None = None, seed: int = 2):
    """PointCloudHierarchyManager::generateHierarchy (PointCloudHierarchy.cc:126-300):
    level 0 the given points, each coarser level the k-means clustering of
    the one below with the weights summed per cluster.  `sizes` overrides the
    reference's level sizes.  Writes the hierarchy files when output_dir is
    given; returns (points per level, densities per level, parent index per
    level below the last)."""
    P = _f64(points)
    dim = P.shape[1] if P.ndim == 2 else None
    if dim is None:
        raise ValueError("points must be [n, dim]")
    if len(P) == 0:
        raise RuntimeError("Input point cloud is empty")
    if densities is None or len(densities) == 0:
        D = np.full(len(P), 1.0 / len(P))
    else:
        D = _f64(densities).ravel()
        if D.size != len(P):
            raise RuntimeError("density vector size doesn't match point cloud size")
    sizes = hierarchy_level_sizes(len(P), min_points, max_points) if sizes is None else list(sizes)
    rng = np.random.default_rng(seed)
    pts, dens, parents = [P], [D], []
    for lev in range(1, len(sizes)):
        k = int(sizes[lev])
        prev_p, prev_d = pts[-1], dens[-1]
        if len(prev_p) <= k:
            c, a = prev_p.copy(), np.arange(len(prev_p))
        else:
            c, a = kmeans(prev_p, k, rng)
        pts.append(c)
        dens.append(np.bincount(a, weights=prev_d, minlength=len(c)))
        parents.append(a.astype(np.int64))
    if output_dir is not None:
        formats.write_target_hierarchy(output_dir, pts, dens, parents)
    return pts, dens, parents


# ---------------------------------------------------------------------------
# drivers
# ---------------------------------------------------------------------------

@dataclass
class MultilevelParameters:
    """SotParameterManager::MultilevelParameters (ParameterManager.h:78-98),
    with its defaults; `subsection multilevel_parameters` of a parameter file
    fills it (parameters.ParameterFile).  The drivers here read the target
    hierarchy fields; the source hierarchy is passed to
    run_combined_multilevel as arrays (mesh generation is not part of this
    library), and the Python clustering switch has no effect (the hierarchy
    is built by generate_target_hierarchy)."""
    source_enabled: bool = False
    source_min_vertices: int = 1000
    source_max_vertices: int = 10000
    source_hierarchy_dir: str = "output/data_multilevel/source_multilevel"
    target_enabled: bool = False
    target_min_points: int = 100
    target_max_points: int = 1000
    target_hierarchy_dir: str = "output/data_multilevel/target_multilevel"
    use_softmax_potential_transfer: bool = True
    use_python_clustering: bool = False
    python_script_name: str = "multilevel_clustering_scipy.py"
    output_prefix: str = "output/multilevel/sot"


@dataclass
class LevelRecord:
    level: int
    n_targets: int
    n_cells: int
    epsilons: list = field(default_factory=list)
    iterations: int = 0
    seconds: float = 0.0
    transfer_seconds: float = 0.0
    threshold: float = 0.0
    status: str = ""


def _copy_params(p: SolverParameters, **kw) -> SolverParameters:
    q = SolverParameters(**{f.name: getattr(p, f.name) for f in fields(p)})
    for k, v in kw.items():
        setattr(q, k, v)
    return q


def _solve_level(solver: SotSolver, psi, params: SolverParameters, level_eps, rec: LevelRecord,
                 save=None):
    """The epsilon loop of one level (SemiDiscreteOT.cc:1153-1238): each
    value solved in turn, the potential carried over; params.epsilon is left
    at the last value, as the reference leaves solver_params.epsilon."""
    for i, eps in enumerate(level_eps):
        params.epsilon = eps
        psi = solver.solve(psi, params)
        info = solver.get_last_solve_info()
        rec.epsilons.append(eps)
        rec.iterations += info["iterations"]
        rec.status = info["status"]
        if save is not None and i < len(level_eps) - 1:
            save(psi, f"potentials_eps{i + 1}")
    return psi


def run_target_multilevel(solver: SotSolver, params: SolverParameters,
                          ml: MultilevelParameters, initial_potential=None,
                          output_dir: str | None = None, io_coding: str = "txt",
                          log=None):
    """SemiDiscreteOT::run_target_multilevel (SemiDiscreteOT.cc:966-1323) on
    the source already set on `solver`: the hierarchy's levels from the
    coarsest, each solved from the potential transferred from the level
    above; with use_epsilon_scaling the epsilon sequence is spread over the
    levels (epsilon_scaling.level_schedules).  Returns (finest potential, records)."""
    files = formats.get_target_hierarchy_files(ml.target_hierarchy_dir)
    if not files:
        raise RuntimeError("No target point cloud hierarchy found")
    n_levels = len(files)
    cur = _copy_params(params)
    original_eps = params.epsilon
    schedules = None
    if params.use_epsilon_scaling:
        schedules = level_schedules(params.epsilon, params.epsilon_scaling_factor,
                                    params.epsilon_scaling_steps, n_levels)
    dim = solver.dim
    level_psi = None
    threshold = 0.0
    y_prev = nu_prev = None
    records = []
    for idx, (pfile, dfile) in enumerate(files):
        lev = formats.level_of(pfile)
        y = formats.read_vector(pfile, dim=dim, file_mode=io_coding)
        nu = formats.read_vector(dfile, file_mode=io_coding)
        rec = LevelRecord(lev, len(nu), solver.n_cells)
        solver.setup_target(y.ravel(), nu)
        if idx > 0:
            t0 = time.perf_counter()
            children = {lev: formats.load_hierarchy_children(ml.target_hierarchy_dir, lev)}
            child_list = [[] for _ in range(lev)] + [children[lev]]
            psi = assign_potentials_by_hierarchy(
                solver, lev + 1, lev, level_psi, y, y_prev, nu_prev, child_list,
                ml.use_softmax_potential_transfer, threshold, cur.epsilon,
                cur.use_log_sum_exp_trick)
            rec.transfer_seconds = time.perf_counter() - t0
        else:
            psi = np.zeros(len(nu)) if initial_potential is None else _f64(initial_potential).copy()
            if psi.size != len(nu):
                raise ValueError("initial potential size does not match the coarsest level")
        out_dir = None if output_dir is None else os.path.join(output_dir, f"target_level_{lev}")
        save = None if out_dir is None else \
            (lambda v, name: formats.write_vector(v, os.path.join(out_dir, name), io_coding))
        t0 = time.perf_counter()
        eps_list = stages_of_level(schedules, idx, original_eps) if schedules else [original_eps]
        psi = _solve_level(solver, psi, cur, eps_list, rec, save)
        rec.seconds = time.perf_counter() - t0
        threshold = solver.get_last_distance_threshold()
        rec.threshold = threshold
        if save is not None:
            save(psi, "potentials")
        records.append(rec)
        if log is not None:
            log(f"target level {lev}: N={len(nu)} eps={rec.epsilons} iterations={rec.iterations} "
                f"({rec.status}) {rec.seconds:.2f}s, transfer {rec.transfer_seconds:.3f}s")
        level_psi, y_prev, nu_prev = psi, y, nu
    if output_dir is not None:
        formats.write_vector(level_psi, os.path.join(output_dir, "potentials"), io_coding)
    return level_psi, records


def run_combined_multilevel(solver: SotSolver, source_levels, params: SolverParameters,
                            ml: MultilevelParameters, initial_potential=None,
                            output_dir: str | None = None, io_coding: str = "txt", log=None):
    """SemiDiscreteOT::run_combined_multilevel (SemiDiscreteOT.cc:1559-1844):
    source levels (coarsest first; each a mesh.SourceArrays) and target
    levels advance together, the coarser list held at its coarsest level
    until the finer list has as many levels left; the potential moves by the
    hierarchy transfer when the target level steps down by one, is copied when
    the target level stays, and the epsilon sequence is spread over the
    combined levels.  Returns (final potential, records)."""
    files = formats.get_target_hierarchy_files(ml.target_hierarchy_dir)
    n_s, n_t = len(source_levels), len(files)
    n_c = max(n_s, n_t)
    cur = _copy_params(params)
    original_eps = params.epsilon
    schedules = None
    if params.use_epsilon_scaling:
        schedules = level_schedules(params.epsilon, params.epsilon_scaling_factor,
                                    params.epsilon_scaling_steps, n_c)
    dim = solver.dim
    cur_psi = None
    prev_t_level, prev_s_idx = -1, -1
    threshold = 0.0
    y_prev = nu_prev = None
    records = []
    for it in range(n_c):
        if n_s >= n_t:
            s_idx = it
            t_idx = it - (n_s - n_t) if it >= n_s - n_t else 0
        else:
            t_idx = it
            s_idx = it - (n_t - n_s) if it >= n_t - n_s else 0
        pfile, dfile = files[t_idx]
        t_level = formats.level_of(pfile)
        if s_idx != prev_s_idx:
            src = source_levels[s_idx]
            solver.setup_source(src.x, src.w, src.centers, src.nq)
            prev_s_idx = s_idx
        y = formats.read_vector(pfile, dim=dim, file_mode=io_coding)
        nu = formats.read_vector(dfile, file_mode=io_coding)
        rec = LevelRecord(t_level, len(nu), solver.n_cells)
        if it == 0:
            psi = np.zeros(len(nu)) if initial_potential is None else _f64(initial_potential).copy()
            if psi.size != len(nu):
                raise ValueError("initial potential size does not match the coarsest level")
        else:
            psi = np.zeros(len(nu))
            if t_level != prev_t_level and prev_t_level == t_level + 1:
                t0 = time.perf_counter()
                child_list = [[] for _ in range(t_level)] + \
                    [formats.load_hierarchy_children(ml.target_hierarchy_dir, t_level)]
                psi = assign_potentials_by_hierarchy(
                    solver, prev_t_level, t_level, cur_psi, y, y_prev, nu_prev, child_list,
                    ml.use_softmax_potential_transfer, threshold, cur.epsilon,
                    cur.use_log_sum_exp_trick)
                rec.transfer_seconds = time.perf_counter() - t0
            elif cur_psi.size == psi.size:
                psi = cur_psi.copy()
        solver.setup_target(y.ravel(), nu)
        out_dir = None if output_dir is None else os.path.join(output_dir, f"combined_level_{it}")
        save = None if out_dir is None else \
            (lambda v, name: formats.write_vector(v, os.path.join(out_dir, name), io_coding))
        t0 = time.perf_counter()
        eps_list = stages_of_level(schedules, it, original_eps) if schedules else [original_eps]
        psi = _solve_level(solver, psi, cur, eps_list, rec, save)
        rec.seconds = time.perf_counter() - t0
        threshold = solver.get_last_distance_threshold()
        rec.threshold = threshold
        if save is not None:
            save(psi, "potentials")
        records.append(rec)
        if log is not None:
            log(f"combined level {it}: cells={rec.n_cells} N={len(nu)} eps={rec.epsilons} "
                f"iterations={rec.iterations} ({rec.status}) {rec.seconds:.2f}s, "
                f"transfer {rec.transfer_seconds:.3f}s")
        cur_psi, prev_t_level, y_prev, nu_prev = psi, t_level, y, nu
    if output_dir is not None and cur_psi is not None:
        formats.write_vector(cur_psi, os.path.join(output_dir, "potentials"), io_coding)
    return cur_psi, records
