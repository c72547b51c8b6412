"""On-disk formats the reference's workflows exchange, read and written on the
host: vectors of doubles or points (Utils::write_vector / read_vector,
include/SemiDiscreteOT/utils/utils.h:55-176, used for potentials by
SemiDiscreteOT::save_results, SemiDiscreteOT.cc:2067-2093) and the target
point-cloud hierarchy (PointCloudHierarchyManager::generateHierarchy,
src/core/PointCloudHierarchy.cc:236-300; read back by
Utils::get_target_hierarchy_files and Utils::load_hierarchy_data,
utils.h:238-396).

Text vectors: one value per line, or one point per line with its coordinates
separated by single spaces (deal.II's Point output), 17 significant digits
(std::numeric_limits<double>::max_digits10).  Binary vectors: the raw
little-endian doubles, points packed coordinate after coordinate.

Hierarchy directory, level 0 the finest:
  level_<k>_points.txt    one point per line
  level_<k>_density.txt   one weight per line
  level_<k>_children.txt  k > 0: per level-k point, "n c_1 ... c_n" (indices into level k-1)
  level_<k>_parents.txt   k < last: per level-k point, "n p_1 ... p_n" (indices into level k+1)

Mesh files as MeshManager::load_source_mesh reads them with deal.II's GridIn
(MeshManager.templates.h:80-140): gmsh 2.2 ASCII (.msh) and legacy VTK ASCII
unstructured grids (.vtk).  The cells of the highest dimension present are
kept (lines and points carry boundary data only); quadrilaterals and
hexahedra are reordered from gmsh / VTK numbering (counter-clockwise faces)
to deal.II's lexicographic vertex order.
"""
from __future__ import annotations

import os
import re

import numpy as np


def _fmt(v: float) -> str:
    return format(float(v), ".17g")


def write_vector(values, filepath: str, file_mode: str = "txt", dim: int | None = None) -> str:
    """Utils::write_vector (utils.h:55-101): `filepath` without extension;
    creates the directory; returns the path written."""
    d = os.path.dirname(filepath)
    if d:
        os.makedirs(d, exist_ok=True)
    a = np.asarray(values, dtype=np.float64)
    if file_mode == "bin":
        path = filepath + ".bin"
        a.astype("<f8").tofile(path)
    elif file_mode == "txt":
        path = filepath + ".txt"
        if dim is None and a.ndim == 2:
            dim = a.shape[1]
        rows = a.reshape(-1, dim) if dim else a.reshape(-1, 1)
        with open(path, "w") as f:
            for r in rows:
                f.write(" ".join(_fmt(v) for v in r) + "\n")
    else:
        raise ValueError(f"Invalid file mode specified: {file_mode}")
    return path


def _resolve(filepath: str, file_mode: str) -> tuple[str, str]:
    """The extension rules of Utils::read_vector (utils.h:108-141)."""
    if not file_mode:
        base, ext = os.path.splitext(filepath)
        if ext in (".txt", ".bin"):
            return filepath, ext[1:]
        return filepath + ".txt", "txt"
    if filepath.endswith("." + file_mode):
        return filepath, file_mode
    return filepath + "." + file_mode, file_mode


def read_vector(filepath: str, dim: int | None = None, file_mode: str = "") -> np.ndarray:
    """Utils::read_vector (utils.h:108-176): every whitespace-separated value
    of a text file, or the raw doubles of a binary one; shaped (n, dim) for
    points.  Raises FileNotFoundError where the reference returns false."""
    path, mode = _resolve(filepath, file_mode)
    if not os.path.exists(path):
        raise FileNotFoundError(f"Unable to open file for reading: {path}")
    if mode == "bin":
        a = np.fromfile(path, dtype="<f8").astype(np.float64)
    else:
        with open(path) as f:
            a = np.array(f.read().split(), dtype=np.float64)
    if dim:
        a = a[: a.size // dim * dim].reshape(-1, dim)
    return a


def get_target_hierarchy_files(directory: str) -> list[tuple[str, str]]:
    """Utils::get_target_hierarchy_files (utils.h:238-275): (points, density)
    paths without extension of every level that has both files, coarsest
    (highest level number) first.  The reference orders by the first digit
    of the level number only, which is the same order for up to 10 levels."""
    if not os.path.isdir(directory):
        raise RuntimeError(f"Target point cloud hierarchy directory does not exist: {directory}")
    out = []
    for name in os.listdir(directory):
        m = re.match(r"^level_(\d+)_points\.txt$", name)
        if not m:
            continue
        k = int(m.group(1))
        dens = os.path.join(directory, f"level_{m.group(1)}_density")
        if os.path.exists(dens + ".txt"):
            out.append((k, os.path.join(directory, f"level_{m.group(1)}_points"), dens))
    out.sort(key=lambda t: -t[0])
    return [(p, d) for _, p, d in out]


def level_of(points_file: str) -> int:
    """The level number the drivers cut out of a points file name
    (SemiDiscreteOT.cc:1084-1088)."""
    name = os.path.basename(points_file)
    return int(name[name.find("level_") + 6: name.find("_points")])


def count_hierarchy_levels(directory: str) -> int:
    """Consecutive level_<k>_points.txt files from k = 0 (utils.h:291-300)."""
    n = 0
    while os.path.exists(os.path.join(directory, f"level_{n}_points.txt")):
        n += 1
    return n


def read_index_lists(path: str) -> list[list[int]]:
    """Lines "n i_1 ... i_n" (utils.h:329-349): at most n indices per line."""
    lists = []
    with open(path) as f:
        for line in f:
            tok = line.split()
            if not tok:
                lists.append([])
                continue
            n = int(tok[0])
            lists.append([int(t) for t in tok[1: 1 + n]])
    return lists


def load_hierarchy_children(directory: str, level: int) -> list[list[int]]:
    """child_indices[level] of Utils::load_hierarchy_data (utils.h:278-396):
    for every point of level+1 its children at `level`, from
    level_<level+1>_children.txt."""
    path = os.path.join(directory, f"level_{level + 1}_children.txt")
    if not os.path.exists(path):
        raise FileNotFoundError(f"Error loading hierarchy data at level {level}")
    return read_index_lists(path)


def write_target_hierarchy(directory: str, points: list, densities: list,
                           parents: list) -> None:
    """The files of PointCloudHierarchyManager::generateHierarchy
    (PointCloudHierarchy.cc:236-300) for levels 0 (finest) .. L-1.
    parents[k] (k < L-1) is the level-(k+1) index of each level-k point."""
    os.makedirs(directory, exist_ok=True)
    L = len(points)
    for k in range(L):
        P = np.asarray(points[k], dtype=np.float64)
        with open(os.path.join(directory, f"level_{k}_points.txt"), "w") as f:
            for r in P:
                f.write(" ".join(_fmt(v) for v in r) + "\n")
        with open(os.path.join(directory, f"level_{k}_density.txt"), "w") as f:
            for v in np.asarray(densities[k], dtype=np.float64):
                f.write(_fmt(v) + "\n")
        if k < L - 1:
            par = np.asarray(parents[k], dtype=np.int64)
            with open(os.path.join(directory, f"level_{k}_parents.txt"), "w") as f:
                for p in par:
                    f.write(f"1 {int(p)}\n")
        if k > 0:
            par = np.asarray(parents[k - 1], dtype=np.int64)
            n_coarse = len(points[k])
            order = np.argsort(par, kind="stable")
            bounds = np.searchsorted(par[order], np.arange(n_coarse + 1))
            with open(os.path.join(directory, f"level_{k}_children.txt"), "w") as f:
                for c in range(n_coarse):
                    ch = order[bounds[c]: bounds[c + 1]]
                    f.write(" ".join([str(len(ch))] + [str(int(i)) for i in ch]) + "\n")


# gmsh element type -> (kind, cell dimension, vertices); VTK cell type likewise
_GMSH = {2: ("simplex", 2, 3), 3: ("q1", 2, 4), 4: ("simplex", 3, 4), 5: ("q1", 3, 8)}
_VTK = {5: ("simplex", 2, 3), 9: ("q1", 2, 4), 10: ("simplex", 3, 4), 12: ("q1", 3, 8)}
# counter-clockwise quad / hex numbering (gmsh and VTK share it) -> lexicographic
_TO_LEX = {4: [0, 1, 3, 2], 8: [0, 1, 3, 2, 4, 5, 7, 6]}


class MeshData:
    """Vertices [n, spacedim] and cells [n_cells, vertices per cell] (lexicographic
    for Q1 cells) of the highest-dimensional cells of a mesh file."""

    def __init__(self, vertices, cells, kind, cell_dim, spacedim):
        self.vertices, self.cells = vertices, cells
        self.kind, self.cell_dim, self.spacedim = kind, cell_dim, spacedim

    @property
    def n_cells(self):
        return len(self.cells)


def _finish(nodes, elems, table, spacedim):
    if not elems:
        raise ValueError("mesh file has no 2D or 3D cells")
    top = max(table[t][1] for t, _ in elems)
    kinds = {table[t][0] for t, _ in elems if table[t][1] == top}
    if len(kinds) != 1:
        raise ValueError("mixed quadrilateral/simplex meshes are not supported")
    kind = kinds.pop()
    cells = np.array([c for t, c in elems if table[t][1] == top], dtype=np.int64)
    if kind == "q1":
        cells = cells[:, _TO_LEX[cells.shape[1]]]
    if spacedim is None:
        spacedim = 3 if (top == 3 or np.any(nodes[:, 2] != 0.0)) else 2
    return MeshData(np.ascontiguousarray(nodes[:, :spacedim]), np.ascontiguousarray(cells), kind,
                    top, spacedim)


def read_msh(path, spacedim=None) -> MeshData:
    """GridIn::read_msh of a gmsh 2.2 ASCII file: $Nodes (tag x y z) and
    $Elements (tag type n_tags tags... vertex tags); node tags may be sparse."""
    with open(path) as f:
        lines = f.read().split("\n")
    it = iter(range(len(lines)))
    nodes, tag_of, elems = None, {}, []
    for i in it:
        line = lines[i].strip()
        if line == "$MeshFormat":
            ver = lines[i + 1].split()
            if not ver or not ver[0].startswith("2") or ver[1] != "0":
                raise ValueError(f"{path}: only gmsh 2.x ASCII files are read (format {ver})")
        elif line == "$Nodes":
            n = int(lines[i + 1])
            rows = np.array([lines[i + 2 + k].split() for k in range(n)], dtype=np.float64)
            tags = rows[:, 0].astype(np.int64)
            tag_of = {int(t): k for k, t in enumerate(tags)}
            nodes = np.zeros((n, 3))
            nodes[:, :rows.shape[1] - 1] = rows[:, 1:4]
        elif line == "$Elements":
            n = int(lines[i + 1])
            for k in range(n):
                v = lines[i + 2 + k].split()
                t = int(v[1])
                if t not in _GMSH:
                    continue
                ntags = int(v[2])
                verts = [tag_of[int(a)] for a in v[3 + ntags:3 + ntags + _GMSH[t][2]]]
                elems.append((t, verts))
    if nodes is None:
        raise ValueError(f"{path}: no $Nodes section")
    return _finish(nodes, elems, _GMSH, spacedim)


def read_vtk(path, spacedim=None) -> MeshData:
    """GridIn::read_vtk of a legacy ASCII UNSTRUCTURED_GRID: POINTS, CELLS and
    CELL_TYPES sections."""
    with open(path) as f:
        tok = f.read().split()
    pos = {w: i for i, w in reversed(list(enumerate(tok))) if w in ("POINTS", "CELLS", "CELL_TYPES")}
    if "POINTS" not in pos or "CELLS" not in pos or "CELL_TYPES" not in pos:
        raise ValueError(f"{path}: not a legacy VTK unstructured grid")
    i = pos["POINTS"]
    n = int(tok[i + 1])
    nodes = np.array(tok[i + 3:i + 3 + 3 * n], dtype=np.float64).reshape(n, 3)
    i = pos["CELLS"]
    nc = int(tok[i + 1])
    k = i + 3
    conn = []
    for _ in range(nc):
        m = int(tok[k])
        conn.append([int(a) for a in tok[k + 1:k + 1 + m]])
        k += 1 + m
    i = pos["CELL_TYPES"]
    types = [int(a) for a in tok[i + 2:i + 2 + nc]]
    elems = [(t, c) for t, c in zip(types, conn) if t in _VTK]
    return _finish(nodes, elems, _VTK, spacedim)


def read_mesh(path, spacedim=None) -> MeshData:
    """By extension, as load_source_mesh dispatches (.vtk, .msh); anything else
    raises "Unsupported file format"."""
    ext = path.rsplit(".", 1)[-1]
    if ext == "msh":
        return read_msh(path, spacedim)
    if ext == "vtk":
        return read_vtk(path, spacedim)
    raise RuntimeError("Unsupported file format: " + ext)


def write_msh(path, vertices, cells, kind, cell_dim):
    """Gmsh 2.2 ASCII with the given cells (lexicographic Q1 order converted
    back to gmsh numbering); the inverse of read_msh."""
    v = np.asarray(vertices, dtype=np.float64)
    if v.shape[1] == 2:
        v = np.hstack([v, np.zeros((len(v), 1))])
    cl = np.asarray(cells, dtype=np.int64)
    gtype = {("simplex", 2): 2, ("q1", 2): 3, ("simplex", 3): 4, ("q1", 3): 5}[(kind, cell_dim)]
    if kind == "q1":
        inv = np.argsort(_TO_LEX[cl.shape[1]])
        cl = cl[:, inv]
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n$Nodes\n%d\n" % len(v))
        for k, p in enumerate(v):
            f.write("%d %s %s %s\n" % (k + 1, _fmt(p[0]), _fmt(p[1]), _fmt(p[2])))
        f.write("$EndNodes\n$Elements\n%d\n" % len(cl))
        for k, c in enumerate(cl):
            f.write("%d %d 2 0 1 %s\n" % (k + 1, gtype, " ".join(str(int(a) + 1) for a in c)))
        f.write("$EndElements\n")
