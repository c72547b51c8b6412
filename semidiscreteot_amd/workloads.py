"""Seeded synthetic workloads (the source meshes, densities and target
measures of the benchmark configurations), built with this package's own
mesh extraction so that the measured path never touches the oracle.

  plain_2d   2D unit square, 64x64 Q1 cells, QGauss(3) (36,864 q), uniform
             density; 1,000 targets U[0,1]^2 (seed 0), nu = 1/1000; eps = 1e-2,
             no truncation, FP64.
  gmm_3d     3D unit cube, 32^3 Q1 hexes, QGauss(2) (262,144 q), density =
             normalized 4-component isotropic Gaussian mixture (means
             U[0.2,0.8]^3, sigma U[0.05,0.15], seed 1); 50,000 targets from a
             6-component mixture (seed 101), nu ~ Dirichlet(5); eps = 1e-3 with
             pointwise truncation.
  ball_tets  Delaunay tetrahedra of 200,000 seeded points in the unit ball
             (seed 3), P1 density 1 + 0.5 sin(4 pi x0), 4-point rule (about
             5.4M q); 500,000 targets uniform in the ball, uniform weights;
             eps = 5e-4 with truncation.
  multilevel_3d  the gmm_3d density on 16^3 -> 32^3 -> 64^3 hexes, targets
             1,000 -> 10,000 -> 100,000 (k-means levels of the 100,000 set,
             seed 2), epsilon 1e-1 -> 1e-2 -> 1e-3.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .mesh import SourceArrays, structured_source


@dataclass
class Workload:
    name: str
    source: SourceArrays
    y: np.ndarray
    nu: np.ndarray
    epsilon: float
    tau: float

    @property
    def dim(self):
        return self.source.dim


def box_nodes(dim, n):
    g = np.linspace(0.0, 1.0, n + 1)
    if dim == 2:
        Y, X = np.meshgrid(g, g, indexing="ij")
        return np.stack([X.ravel(), Y.ravel()], 1)
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)


def gmm_pdf(points, means, sig):
    dim = points.shape[1]
    val = np.zeros(len(points))
    for m, s in zip(means, sig):
        val += np.exp(-0.5 * np.sum((points - m) ** 2, axis=1) / s ** 2) / (2 * np.pi * s * s) ** (dim / 2)
    return val / len(means)


def gmm_points(rng, n, dim, k, alpha=1.0):
    means = rng.uniform(0.2, 0.8, size=(k, dim))
    sig = rng.uniform(0.05, 0.15, size=k)
    wts = rng.dirichlet(np.full(k, alpha))
    comp = rng.choice(k, size=n, p=wts)
    return means[comp] + sig[comp, None] * rng.standard_normal((n, dim))


def plain_2d(n_cells=64, n_targets=1000, seed=0) -> Workload:
    src = structured_source(2, [n_cells, n_cells], [0.0, 0.0], [1.0, 1.0], 3)
    rng = np.random.default_rng(seed)
    y = rng.random((n_targets, 2))
    nu = np.full(n_targets, 1.0 / n_targets)
    return Workload("plain_2d_64x64_q3_N1000", src, y.ravel(), nu, 1e-2, 0.0)


def gmm_3d(n_cells=32, n_targets=50000, seed=1, eps=1e-3, tau=1e-16) -> Workload:
    rng = np.random.default_rng(seed)
    means = rng.uniform(0.2, 0.8, size=(4, 3))
    sig = rng.uniform(0.05, 0.15, size=4)
    rho = gmm_pdf(box_nodes(3, n_cells), means, sig)
    src = structured_source(3, [n_cells] * 3, [0.0] * 3, [1.0] * 3, 2, rho)
    trng = np.random.default_rng(seed + 100)
    y = gmm_points(trng, n_targets, 3, 6)
    nu = trng.dirichlet(np.full(n_targets, 5.0))
    nu = nu / nu.sum()
    return Workload(f"gmm_3d_{n_cells}cube_q2_N{n_targets}", src, y.ravel(), nu, eps, tau)


def ball_points(rng, n):
    """n points uniform in the unit ball."""
    p = rng.standard_normal((n, 3))
    p *= (rng.random(n) ** (1.0 / 3.0) / np.linalg.norm(p, axis=1))[:, None]
    return p


def ball_tets(n_points=200000, n_targets=500000, seed=3, eps=5e-4, tau=1e-12) -> Workload:
    """Large unstructured workload: Delaunay tetrahedra (scipy) of n_points
    seeded points uniform in the unit ball (seed 3), P1 density
    1 + 0.5 sin(4 pi x0) at the vertices, QGaussSimplex(2) (4 points per
    tet); n_targets targets uniform in the ball, uniform weights; eps = 5e-4
    with pointwise truncation.  Zero-volume slivers of the triangulation are
    dropped (deal.II rejects them)."""
    from scipy.spatial import Delaunay
    from .mesh import simplex_source
    rng = np.random.default_rng(seed)
    p = ball_points(rng, n_points)
    cells = Delaunay(p).simplices.astype(np.int64)
    v = p[cells]
    vol = np.abs(np.einsum("ij,ij->i", v[:, 1] - v[:, 0],
                           np.cross(v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]))) / 6.0
    cells = np.ascontiguousarray(cells[vol > 1e-12])
    rho = 1.0 + 0.5 * np.sin(4.0 * np.pi * p[:, 0])
    src = simplex_source(3, p.ravel(), cells.ravel(), 2, rho)
    y = ball_points(rng, n_targets)
    nu = np.full(n_targets, 1.0 / n_targets)
    return Workload(f"ball_tets_{n_points}pts_p1q4_N{n_targets}", src, y.ravel(), nu, eps, tau)


@dataclass
class MultilevelWorkload:
    name: str
    sources: list          # SourceArrays per source level, coarsest first
    hierarchy: tuple       # (points, densities, parents) per target level, level 0 finest
    epsilon: float         # finest epsilon
    epsilon_factor: float
    epsilon_steps: int
    tau: float


def multilevel_3d(cells=(16, 32, 64), n_targets=100000, sizes=(100000, 10000, 1000),
                  seed=1, target_seed=101, kmeans_seed=2, eps=1e-3, eps_factor=10.0,
                  tau=1e-12) -> MultilevelWorkload:
    """Combined multilevel workload: the gmm_3d source density on 16^3, 32^3
    and 64^3 Q1 hexes with QGauss(2) (32,768 / 262,144 / 2,097,152 q), the
    gmm_3d target mixture with 100,000 points (nu ~ Dirichlet(5)) and its
    coarse levels of 10,000 and 1,000 points from seeded k-means (seed 2)
    with the weights summed per cluster; epsilon annealed 1e-1 -> 1e-2 -> 1e-3
    (one value per combined level)."""
    from .multilevel import generate_target_hierarchy
    rng = np.random.default_rng(seed)
    means = rng.uniform(0.2, 0.8, size=(4, 3))
    sig = rng.uniform(0.05, 0.15, size=4)
    sources = []
    for n in cells:
        rho = gmm_pdf(box_nodes(3, n), means, sig)
        sources.append(structured_source(3, [n] * 3, [0.0] * 3, [1.0] * 3, 2, rho))
    trng = np.random.default_rng(target_seed)
    y = gmm_points(trng, n_targets, 3, 6)
    nu = trng.dirichlet(np.full(n_targets, 5.0))
    nu = nu / nu.sum()
    hier = generate_target_hierarchy(y, nu, None, sizes=list(sizes), seed=kmeans_seed)
    name = "multilevel_3d_" + "-".join(f"{n}cube" for n in cells) + "_N" + \
        "-".join(str(s) for s in reversed(sizes))
    return MultilevelWorkload(name, sources, hier, eps, eps_factor, len(cells), tau)


@dataclass
class Request:
    index: int
    arrival: float       # seconds after the start of the trace
    y: np.ndarray        # [N*3]
    nu: np.ndarray       # [N]
    epsilon: float
    tolerance: float


def serving_trace(n_requests=10000, rate=50.0, seed=4, n_min=256, n_max=65536,
                  eps_choices=(1e-2, 3e-3, 1e-3), tolerance=1e-6):
    """Request stream of the serving workload: Poisson arrivals at `rate`
    requests/s (seed 4); each request has N ~ log-uniform [256, 65536] targets
    from a seeded 3-8 component Gaussian mixture, nu ~ Dirichlet(2), eps drawn
    from {1e-2, 3e-3, 1e-3}, componentwise tolerance 1e-6.  Requests are
    generated lazily (request i uses its own seeded stream)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n_requests)
    arrivals = np.cumsum(gaps) - gaps[0]
    for i in range(n_requests):
        r = np.random.default_rng([seed, i])
        N = int(np.exp(r.uniform(np.log(n_min), np.log(n_max))))
        k = int(r.integers(3, 9))
        y = gmm_points(r, N, 3, k)
        nu = r.dirichlet(np.full(N, 2.0))
        nu = nu / nu.sum()
        eps = float(eps_choices[int(r.integers(0, len(eps_choices)))])
        yield Request(i, float(arrivals[i]), y.ravel(), nu, eps, tolerance)
