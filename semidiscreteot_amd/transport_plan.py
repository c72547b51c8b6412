"""Transport maps after a solve: the reference's OptimalTransportPlan<spacedim>
(include/SemiDiscreteOT/core/OptimalTransportPlan.h, src/core/
OptimalTransportPlan.cc:8-96) with its modal and barycentric strategies
(:99-189, :218-310), on the GPU (sdot_transport_map, FP64).

    plan = OptimalTransportPlan(3, "barycentric")
    plan.set_source_measure(x, rho); plan.set_target_measure(y, nu)
    plan.set_potential(psi, eps); plan.set_truncation_radius(r)
    plan.compute_map(); plan.mapped_points, plan.transport_density

The source measure is a list of points with densities, as in the reference;
the points are held by a context as one-point cells (the map needs only the
target layout).  Unknown strategy names raise, as create_strategy does.
"""
from __future__ import annotations

import os

import numpy as np

from . import _lib as L
from .formats import write_vector
from .solver import SotSolver, _f64, _p

_STRATEGIES = {"modal": 0, "barycentric": 1}


class OptimalTransportPlan:
    def __init__(self, dim: int, strategy_name: str = "modal", device: int = 0):
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        self.dim = dim
        self.device = device
        self.set_strategy(strategy_name)
        self.source_points = None
        self.source_density = None
        self.target_points = None
        self.target_density = None
        self.transport_potential = None
        self.epsilon = 1e-2
        self.truncation_radius = -1.0
        self.mapped_points = None
        self.transport_density = None

    @staticmethod
    def get_available_strategies():
        return list(_STRATEGIES)

    def set_strategy(self, strategy_name: str):
        if strategy_name not in _STRATEGIES:
            raise RuntimeError("Unknown strategy: " + strategy_name)
        self.strategy = strategy_name

    def set_source_measure(self, points, density):
        pts = _f64(points).reshape(-1, self.dim)
        rho = _f64(density).ravel()
        if len(pts) != rho.size:
            raise ValueError("points and density differ in size")
        self.source_points, self.source_density = pts, rho

    def set_target_measure(self, points, density):
        pts = _f64(points).reshape(-1, self.dim)
        nu = _f64(density).ravel()
        if len(pts) != nu.size:
            raise ValueError("points and density differ in size")
        self.target_points, self.target_density = pts, nu

    def set_potential(self, potential, regularization_param: float):
        psi = _f64(potential).ravel()
        if self.target_points is None or psi.size != len(self.target_points):
            raise ValueError("potential size differs from the target count")
        self.transport_potential = psi
        self.epsilon = float(regularization_param)

    def set_truncation_radius(self, radius: float):
        self.truncation_radius = float(radius)

    def compute_map(self):
        if self.source_points is None or len(self.source_points) == 0:
            raise RuntimeError("Source measure not set")
        if self.target_points is None or len(self.target_points) == 0:
            raise RuntimeError("Target measure not set")
        if self.transport_potential is None:
            raise RuntimeError("Transport potential not set")
        x = self.source_points
        with SotSolver(self.dim, device=self.device) as s:
            # one-point cells: the context needs a source before its targets
            s.setup_source(x.ravel(), np.abs(self.source_density), x.ravel(), 1)
            s.setup_target(self.target_points.ravel(), self.target_density)
            out = np.zeros(len(x) * self.dim)
            L.check(s.lib.sdot_transport_map(s._h, _STRATEGIES[self.strategy],
                                             _p(self.transport_potential), self.epsilon,
                                             self.truncation_radius, len(x), _p(_f64(x.ravel())),
                                             _p(out)))
        self.mapped_points = out.reshape(-1, self.dim)
        self.transport_density = self.source_density.copy()
        return self.mapped_points

    def save_map(self, output_dir: str):
        """The text files of save_results (OptimalTransportPlan.cc:192-196):
        mapped_points.txt and transport_density.txt (the VTK files are not
        written)."""
        if self.mapped_points is None:
            raise RuntimeError("compute_map first")
        os.makedirs(output_dir, exist_ok=True)
        write_vector(self.mapped_points, os.path.join(output_dir, "mapped_points"), "txt")
        write_vector(self.transport_density, os.path.join(output_dir, "transport_density"), "txt")
