"""Lloyd's algorithm over semi-discrete OT (Lloyd<dim>::run_lloyd,
src/core/Lloyd.cc:14-60, with run_sot_iteration :180-281 and
run_centroid_iteration :82-177): alternate a solve for the potential (from
zero each time, with the epsilon sequence when scaling is on) and the move of
every target to its weighted barycenter, until the root-mean-square step is
below the threshold.  Each solve and barycenter evaluation runs on the GPU.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, fields

import numpy as np

from .epsilon_scaling import level_schedules
from .solver import SolverParameters, SotSolver, _f64


@dataclass
class LloydRecord:
    iteration: int
    solve_iterations: int
    status: str
    step_norm: float


def run_lloyd(solver: SotSolver, target_points, target_density, params: SolverParameters,
              absolute_threshold: float = 1e-8, max_iterations: int = 100, log=None):
    """Returns (barycenters [N, dim], potential, records).  `solver` holds the
    source measure; the targets start at target_points."""
    dim = solver.dim
    nu = _f64(target_density).ravel()
    bary = _f64(target_points).reshape(-1, dim).copy()
    prev = bary.copy()
    solver.setup_target(bary.ravel(), nu)
    cur = SolverParameters(**{f.name: getattr(params, f.name) for f in fields(params)})
    seq = [params.epsilon]
    if params.use_epsilon_scaling:
        seq = level_schedules(params.epsilon, params.epsilon_scaling_factor,
                              params.epsilon_scaling_steps, 1)[0]
    records = []
    psi = None
    for n_iter in range(max_iterations):
        # run_sot_iteration: the potential restarts from zero (Lloyd.cc:198)
        psi = np.zeros(nu.size)
        its, status = 0, ""
        for eps in seq:
            cur.epsilon = eps
            psi = solver.solve(psi, cur)
            info = solver.get_last_solve_info()
            its += info["iterations"]
            status = info["status"]
        # run_centroid_iteration at the last solve's epsilon
        bary = solver.evaluate_weighted_barycenters(psi, cur)
        # compute_step_norm (Lloyd.cc:64-80): root-mean-square move
        l2 = math.sqrt(float(np.sum((bary - prev) ** 2)) / len(bary))
        prev = bary.copy()
        records.append(LloydRecord(n_iter, its, status, l2))
        if log is not None:
            log(f"lloyd iteration {n_iter + 1}: solve {its} iterations ({status}), step {l2:.3e}")
        if l2 < absolute_threshold:
            break
        solver.setup_target(bary.ravel(), nu)
    return bary, psi, records
