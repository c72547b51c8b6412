"""Mesh-file ingestion and general-cell source extraction (CPU): the reading
MeshManager::load_source_mesh does with deal.II's GridIn
(MeshManager.templates.h:80-140) and the FEValues values of general Q1 / P1
cells, also as surfaces in 3D.

The oracle's general-cell restatement (orc_cells_source) is pinned by closed
forms: affine boxes reproduce the structured builder, a hexahedron with one
raised vertex has volume 1 + s/4 and first z-moment (1 + s/2 + s^2/9)/2, a
tilted planar quadrilateral in 3D has area sqrt(1 + h^2), a bilinear saddle
quadrilateral's area matches a 40-digit mpmath quadrature, and a surface
triangle has area |e1 x e2|/2.  The product builder (sdot_mesh_cells) must
agree with the oracle to 1e-14 relative on distorted meshes."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from semidiscreteot_amd import formats
from semidiscreteot_amd.mesh import cells_source, source_from_mesh_file
from tests.cases import box_nodes

REF = "/root/reference/examples"


def lex_cells(dim, n):
    """Cells of the n^dim box grid in deal.II's lexicographic vertex order."""
    out = []
    nv = n + 1
    rng = range(n)
    if dim == 2:
        for j in rng:
            for i in rng:
                b = i + nv * j
                out.append([b, b + 1, b + nv, b + nv + 1])
    else:
        for k in rng:
            for j in rng:
                for i in rng:
                    b = i + nv * (j + nv * k)
                    out.append([b, b + 1, b + nv, b + nv + 1,
                                b + nv * nv, b + nv * nv + 1, b + nv * nv + nv, b + nv * nv + nv + 1])
    return np.array(out, dtype=np.int64)


@pytest.mark.parametrize("dim,qorder", [(2, 2), (2, 3), (3, 2)])
def test_oracle_cells_affine_box_equals_structured(dim, qorder):
    n = 5
    v = box_nodes(dim, n)
    rho = 1.0 + v[:, 0] + 0.5 * v[:, 1] ** 2
    a = O.cells_source(True, dim, dim, v, lex_cells(dim, n), qorder, rho)
    b = O.structured_source(dim, [n] * dim, [0.0] * dim, [1.0] * dim, qorder, rho)
    for u, w in ((a.x, b.x), (a.jxw, b.jxw), (a.rho, b.rho), (a.center, b.center)):
        assert np.max(np.abs(u - w)) <= 1e-15 * max(1.0, np.max(np.abs(w)))


def test_oracle_cells_closed_forms():
    s = 0.3
    v = box_nodes(3, 1).copy()
    v[7, 2] += s   # vertex (1,1,1) -> (1,1,1+s): z = xi3 (1 + s xi1 xi2)
    src = O.cells_source(True, 3, 3, v, lex_cells(3, 1), 2, normalize=False)
    assert abs(src.jxw.sum() - (1 + s / 4)) <= 1e-15
    z = src.x.reshape(-1, 3)[:, 2]
    assert abs(np.sum(src.jxw * z) - 0.5 * (1 + s / 2 + s * s / 9)) <= 1e-15
    # tilted planar quadrilateral in 3D
    h = 0.75
    q = np.array([[0, 0, 0], [1, 0, 0], [0, 1, h], [1, 1, h]], dtype=float)
    t = O.cells_source(True, 2, 3, q, np.array([[0, 1, 2, 3]]), 1, normalize=False)
    assert abs(t.jxw.sum() - np.sqrt(1 + h * h)) <= 1e-15
    # bilinear saddle z = 0.7 xi1 xi2: area from a 40-digit mpmath quadrature
    q = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.7]], dtype=float)
    t = O.cells_source(True, 2, 3, q, np.array([[0, 1, 2, 3]]), 10, normalize=False)
    assert abs(t.jxw.sum() - 1.14837975588428145985489958516) <= 1e-13
    # surface triangle (0,0,0), (1,0,0), (0,1,1): area sqrt(2)/2
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 1]], dtype=float)
    t = O.cells_source(False, 2, 3, tri, np.array([[0, 1, 2]]), 2, normalize=False)
    assert abs(t.jxw.sum() - np.sqrt(2) / 2) <= 1e-15


def test_oracle_cells_simplex_equals_simplex_source():
    rng = np.random.default_rng(3)
    v = rng.random((30, 3))
    from scipy.spatial import Delaunay
    cells = Delaunay(v).simplices.astype(np.int64)
    rho = 1 + rng.random(30)
    a = O.cells_source(False, 3, 3, v, cells, 2, rho)
    b = O.simplex_source(3, v.ravel(), cells.ravel(), 2, rho)
    for u, w in ((a.x, b.x), (a.jxw, b.jxw), (a.rho, b.rho), (a.center, b.center)):
        assert np.max(np.abs(u - w)) <= 1e-15 * max(1.0, np.max(np.abs(w)))


def _distorted(dim_c, dim_s, n, seed):
    v = box_nodes(dim_c, n)
    rng = np.random.default_rng(seed)
    v = v + 0.15 / n * (rng.random(v.shape) - 0.5)
    if dim_s > dim_c:
        # the 2D grid bent onto a cylinder patch in 3D
        th = v[:, 0] * 1.2
        v = np.stack([np.cos(th), np.sin(th), v[:, 1]], 1)
    return v, lex_cells(dim_c, n), 1.0 + rng.random(len(v))


@pytest.mark.parametrize("cdim,sdim,qorder", [(2, 2, 3), (3, 3, 2), (2, 3, 2)])
def test_product_cells_match_oracle(cdim, sdim, qorder):
    v, cells, rho = _distorted(cdim, sdim, 4, cdim + sdim)
    a = cells_source("q1", cdim, sdim, v, cells, qorder, rho)
    b = O.cells_source(True, cdim, sdim, v, cells, qorder, rho)
    assert a.dim == sdim and a.nq == b.nq
    assert np.max(np.abs(a.x - b.x)) <= 1e-14
    assert np.max(np.abs(a.w - b.w)) <= 1e-14 * np.max(b.w)
    assert np.max(np.abs(a.centers - b.center)) <= 1e-15
    assert abs(a.w.sum() - 1.0) <= 1e-13   # normalised density


def test_product_surface_triangles_match_oracle():
    rng = np.random.default_rng(8)
    th = rng.random(40) * 2 * np.pi
    z = rng.random(40)
    v = np.stack([np.cos(th), np.sin(th), z], 1)
    from scipy.spatial import Delaunay
    cells = Delaunay(np.stack([th, z], 1)).simplices.astype(np.int64)
    a = cells_source("simplex", 2, 3, v, cells, 2)
    b = O.cells_source(False, 2, 3, v, cells, 2)
    assert np.max(np.abs(a.x - b.x)) <= 1e-15 and np.max(np.abs(a.w - b.w)) <= 1e-14 * np.max(b.w)


@pytest.mark.parametrize("kind,cdim", [("q1", 3), ("q1", 2), ("simplex", 3)])
def test_msh_round_trip(tmp_path, kind, cdim):
    if kind == "q1":
        v, cells, rho = _distorted(cdim, cdim, 3, 11)
    else:
        rng = np.random.default_rng(4)
        v = rng.random((25, 3))
        from scipy.spatial import Delaunay
        cells = Delaunay(v).simplices.astype(np.int64)
        rho = None
    path = str(tmp_path / "mesh.msh")
    formats.write_msh(path, v, cells, kind, cdim)
    m = formats.read_msh(path)
    assert (m.kind, m.cell_dim, m.spacedim) == (kind, cdim, cdim)
    assert np.array_equal(m.cells, cells) and np.array_equal(m.vertices, v)
    a, _ = source_from_mesh_file(path, 2, rho)
    b = cells_source(kind, cdim, cdim, v, cells, 2, rho)
    assert np.array_equal(a.x, b.x) and np.array_equal(a.w, b.w)


def test_vtk_reader(tmp_path):
    """A legacy VTK file with one hexahedron (VTK numbering) and a boundary
    quad: the hex is kept, reordered lexicographically."""
    pts = box_nodes(3, 1)                       # lexicographic
    vtk_order = [0, 1, 3, 2, 4, 5, 7, 6]        # lexicographic index of VTK vertex k
    text = ["# vtk DataFile Version 3.0", "cube", "ASCII", "DATASET UNSTRUCTURED_GRID",
            "POINTS 8 double"] + [" ".join(map(str, p)) for p in pts] + \
        ["CELLS 2 15", "8 " + " ".join(map(str, vtk_order)), "4 0 1 3 2",
         "CELL_TYPES 2", "12", "9"]
    path = tmp_path / "cube.vtk"
    path.write_text("\n".join(text) + "\n")
    m = formats.read_mesh(str(path))
    assert (m.kind, m.cell_dim, m.spacedim, m.n_cells) == ("q1", 3, 3, 1)
    assert m.cells.tolist() == [list(range(8))]
    with pytest.raises(RuntimeError):
        formats.read_mesh(str(tmp_path / "cube.stl"))


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference examples not present")
def test_reference_example_meshes():
    """The reference's own example meshes: tutorial_0's tetrahedral source and
    tutorial_4's quadrilateral surfaces of a sphere and a cylinder."""
    src, m = source_from_mesh_file(os.path.join(REF, "tutorial_0/run/source.msh"), 2)
    assert (m.kind, m.cell_dim, m.spacedim, m.n_cells) == ("simplex", 3, 3, 98304)
    assert src.nq == 4 and abs(src.w.sum() - 1.0) <= 1e-10   # 393,216 summed terms
    raw, _ = source_from_mesh_file(os.path.join(REF, "tutorial_0/run/source.msh"), 2,
                                   normalize=False)
    vol = raw.w.sum()
    lo, hi = m.vertices.min(0), m.vertices.max(0)
    assert 0 < vol <= np.prod(hi - lo) * (1 + 1e-12)
    sph, m = source_from_mesh_file(os.path.join(REF, "tutorial_4/run/sphere.msh"), 2,
                                   normalize=False)
    assert (m.kind, m.cell_dim, m.spacedim) == ("q1", 2, 3)
    r = np.linalg.norm(m.vertices, axis=1).mean()
    assert abs(sph.w.sum() / (4 * np.pi * r * r) - 1.0) < 0.02   # polyhedral approximation
