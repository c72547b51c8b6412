"""Source-measure extraction (the FEValues stand-in) over the C ABI:
quadrature points, rho*JxW and cell centers of structured Q1 boxes, P1
simplex meshes and general Q1 / P1 cells (also surfaces in 3D) read from mesh
files (utils.h:776-840 element/quadrature choice, density normalisation of
SemiDiscreteOT.cc:422-449, GridIn reading of MeshManager.templates.h:80-140)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L


@dataclass
class SourceArrays:
    dim: int
    nq: int
    x: np.ndarray        # [n_cells*nq*dim]
    w: np.ndarray        # [n_cells*nq]  rho * JxW
    centers: np.ndarray  # [n_cells*dim]

    @property
    def n_cells(self):
        return self.centers.size // self.dim


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def structured_source(dim, n_sub, lo, hi, qorder, nodal_rho=None, normalize=True) -> SourceArrays:
    ns = [int(v) for v in n_sub] + [1] * (3 - len(n_sub))
    n_cells = int(np.prod(ns[:dim]))
    nq = int(qorder) ** dim
    x = np.zeros(n_cells * nq * dim)
    w = np.zeros(n_cells * nq)
    c = np.zeros(n_cells * dim)
    lo = np.ascontiguousarray(lo, dtype=np.float64)
    hi = np.ascontiguousarray(hi, dtype=np.float64)
    rho = None if nodal_rho is None else np.ascontiguousarray(nodal_rho, dtype=np.float64)
    if rho is not None:
        assert rho.size == int(np.prod([n + 1 for n in ns[:dim]]))
    L.check(L.load().sdot_mesh_structured(dim, (C.c_int * 3)(*ns), _p(lo), _p(hi), int(qorder),
                                          None if rho is None else _p(rho), int(bool(normalize)),
                                          _p(x), _p(w), _p(c)))
    return SourceArrays(dim, nq, x, w, c)


def simplex_source(dim, vertices, cells, qorder, nodal_rho=None, normalize=True) -> SourceArrays:
    v = np.ascontiguousarray(vertices, dtype=np.float64)
    cl = np.ascontiguousarray(cells, dtype=np.int64)
    n_cells = cl.size // (dim + 1)
    nq = {1: 1, 2: dim + 1}.get(int(qorder))
    if nq is None:
        raise L.SdotError(-6, f"QGaussSimplex order {qorder} is not provided (orders 1 and 2)")
    x = np.zeros(n_cells * nq * dim)
    w = np.zeros(n_cells * nq)
    c = np.zeros(n_cells * dim)
    rho = None if nodal_rho is None else np.ascontiguousarray(nodal_rho, dtype=np.float64)
    L.check(L.load().sdot_mesh_simplex(dim, v.size // dim, _p(v), n_cells,
                                       cl.ctypes.data_as(C.POINTER(C.c_int64)), int(qorder),
                                       None if rho is None else _p(rho), int(bool(normalize)),
                                       _p(x), _p(w), _p(c)))
    return SourceArrays(dim, nq, x, w, c)


def cells_source(kind, cell_dim, spacedim, vertices, cells, qorder, nodal_rho=None,
                 normalize=True) -> SourceArrays:
    """General cells: kind "q1" (2^cell_dim vertices per cell, deal.II's
    lexicographic order; QGauss(qorder)) or "simplex" (QGaussSimplex(qorder),
    orders 1-2), cells of dimension cell_dim in a space of dimension spacedim
    (surfaces in 3D: cell_dim 2, spacedim 3)."""
    k = {"simplex": 0, "q1": 1}[kind]
    v = np.ascontiguousarray(vertices, dtype=np.float64)
    cl = np.ascontiguousarray(cells, dtype=np.int64)
    nv = 2 ** cell_dim if k == 1 else cell_dim + 1
    n_cells = cl.size // nv
    if k == 1:
        nq = int(qorder) ** cell_dim
    else:
        nq = {1: 1, 2: cell_dim + 1}.get(int(qorder))
        if nq is None:
            raise L.SdotError(-6, f"QGaussSimplex order {qorder} is not provided (orders 1 and 2)")
    x = np.zeros(n_cells * nq * spacedim)
    w = np.zeros(n_cells * nq)
    c = np.zeros(n_cells * spacedim)
    rho = None if nodal_rho is None else np.ascontiguousarray(nodal_rho, dtype=np.float64)
    L.check(L.load().sdot_mesh_cells(k, int(cell_dim), int(spacedim), v.size // spacedim, _p(v),
                                     n_cells, cl.ctypes.data_as(C.POINTER(C.c_int64)), int(qorder),
                                     None if rho is None else _p(rho), int(bool(normalize)),
                                     _p(x), _p(w), _p(c)))
    return SourceArrays(spacedim, nq, x, w, c)


def source_from_mesh_file(path, qorder, nodal_rho=None, normalize=True, spacedim=None):
    """The source measure of a mesh file as MeshManager::load_source_mesh reads
    it (.msh, gmsh 2.2 ASCII; .vtk, legacy ASCII unstructured grid) with the
    element / quadrature choice of Utils::create_quadrature_for_mesh: Q1
    cells with QGauss(qorder), simplices with QGaussSimplex(qorder).  Returns
    (SourceArrays, MeshData)."""
    from .formats import read_mesh
    m = read_mesh(path, spacedim=spacedim)
    src = cells_source(m.kind, m.cell_dim, m.spacedim, m.vertices, m.cells, qorder, nodal_rho,
                       normalize)
    return src, m
