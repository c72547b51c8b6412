"""Seeded inputs shared by the CPU and GPU tests (and by bench.py's parity
cases): sources built with the oracle's deal.II stand-in, targets from
uniform or Gaussian-mixture samplers."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O


def gmm_sample(rng, n, dim, k, lo=0.2, hi=0.8, smin=0.05, smax=0.15, box=(0.0, 1.0)):
    means = rng.uniform(lo, hi, size=(k, dim))
    sig = rng.uniform(smin, smax, size=k)
    wts = rng.dirichlet(np.ones(k))
    comp = rng.choice(k, size=n, p=wts)
    pts = means[comp] + sig[comp, None] * rng.standard_normal((n, dim))
    return np.clip(pts, box[0], box[1])


def gmm_density(points, rng, dim, k):
    means = rng.uniform(0.2, 0.8, size=(k, dim))
    sig = rng.uniform(0.05, 0.15, size=k)
    val = np.zeros(len(points))
    for m, s in zip(means, sig):
        val += np.exp(-0.5 * np.sum((points - m) ** 2, axis=1) / s ** 2) / (2 * np.pi * s * s) ** (dim / 2)
    return val / k


def box_nodes(dim, n):
    g = np.linspace(0.0, 1.0, n + 1)
    if dim == 2:
        X, Y = np.meshgrid(g, g, indexing="xy")
        return np.stack([X.ravel(), Y.ravel()], 1)
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)


def structured_case(dim, n, qorder, N, seed, gmm_source=False, gmm_targets=False,
                    dirichlet=None):
    rng = np.random.default_rng(seed)
    nodal = None
    if gmm_source:
        nodal = gmm_density(box_nodes(dim, n), rng, dim, 4)
    src = O.structured_source(dim, [n] * dim, [0.0] * dim, [1.0] * dim, qorder, nodal)
    if gmm_targets:
        y = gmm_sample(rng, N, dim, 6)
    else:
        y = rng.random((N, dim))
    if dirichlet is not None:
        nu = rng.dirichlet(np.full(N, float(dirichlet)))
    else:
        nu = np.full(N, 1.0 / N)
    nu = nu / nu.sum()
    return src, O.Target(y.ravel(), nu)


def tet_case(n_points, N, seed):
    """Delaunay tets of seeded points in the unit ball (scipy), P1, 4-point rule,
    density 1 + 0.5 sin(4 pi x0)."""
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    p = rng.standard_normal((n_points, 3))
    p *= (rng.random(n_points) ** (1 / 3) / np.linalg.norm(p, axis=1))[:, None]
    tri = Delaunay(p)
    cells = tri.simplices.astype(np.int64)
    # drop degenerate slivers (deal.II would reject zero-volume cells)
    v = p[cells]
    vol = np.abs(np.einsum("ij,ij->i", v[:, 1] - v[:, 0], np.cross(v[:, 2] - v[:, 0], v[:, 3] - v[:, 0]))) / 6
    cells = cells[vol > 1e-12]
    rho = 1.0 + 0.5 * np.sin(4 * np.pi * p[:, 0])
    src = O.simplex_source(3, p.ravel(), cells.ravel(), 2, rho)
    q = rng.standard_normal((N, 3))
    q *= (rng.random(N) ** (1 / 3) / np.linalg.norm(q, axis=1))[:, None]
    return src, O.Target(q.ravel(), np.full(N, 1.0 / N)), (p, cells, rho)
