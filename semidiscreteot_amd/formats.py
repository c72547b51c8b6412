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
