"""Host-side pieces of the multilevel callers (no GPU): the epsilon
schedule and its split over levels (EpsilonScalingHandler.cc:6-100), the vector and
hierarchy file formats (utils.h:55-176, 238-396; PointCloudHierarchy.cc:
236-300), the hierarchy's level sizes and clustering, and the direct
coarse-to-fine potential assignment (SemiDiscreteOT.cc:808-826)."""
import os

import numpy as np
import pytest

from semidiscreteot_amd import formats
from semidiscreteot_amd.epsilon_scaling import (epsilon_schedule, level_schedules,
                                                normalized_scaling, split_schedule,
                                                stages_of_level)
from semidiscreteot_amd import multilevel as ml


def test_epsilon_sequence_and_fallbacks():
    # stage i at eps * factor^(n-1-i); factor <= 1 -> 2, 0 stages -> 1
    assert epsilon_schedule(1e-3, 10.0, 3) == [1e-3 * 10.0 ** 2, 1e-3 * 10.0, 1e-3]
    assert epsilon_schedule(1e-2, 1.0, 2) == [2e-2, 1e-2]
    assert epsilon_schedule(1e-2, 0.5, 3) == [4e-2, 2e-2, 1e-2]
    assert epsilon_schedule(1e-2, 3.0, 0) == [1e-2]
    assert normalized_scaling(1.0, 0) == (2.0, 1)


def test_epsilon_distribution_over_levels():
    seq = epsilon_schedule(1.0, 2.0, 5)                  # 16, 8, 4, 2, 1
    assert seq == [16.0, 8.0, 4.0, 2.0, 1.0]
    # more values than levels: ceil(values left / levels left), coarsest first
    assert split_schedule(seq, 3) == [seq[0:2], seq[2:4], seq[4:5]]
    assert split_schedule(seq, 2) == [seq[0:3], seq[3:5]]
    assert level_schedules(1.0, 2.0, 7, 3) == [[64.0, 32.0, 16.0], [8.0, 4.0], [2.0, 1.0]]
    # every value is used exactly once, in order
    for n in range(1, 9):
        for L in range(1, 9):
            parts = split_schedule(list(range(n)), L)
            assert len(parts) == L
            if n > L:
                assert sum(parts, []) == list(range(n))
                assert all(len(a) >= len(b) for a, b in zip(parts, parts[1:]))
    # fewer values than levels: one each, the smallest repeated
    assert level_schedules(1.0, 2.0, 2, 4) == [[2.0], [1.0], [1.0], [1.0]]
    assert split_schedule([3.0], 0) == []
    assert stages_of_level([[2.0], [1.0]], 7, 5.0) == [5.0]
    assert stages_of_level([[2.0], [1.0]], 1, 5.0) == [1.0]


def test_vector_formats_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    v = rng.standard_normal(257) * 10.0 ** rng.integers(-30, 30, 257)
    p = formats.write_vector(v, str(tmp_path / "sub" / "potentials"))
    assert p.endswith("potentials.txt")
    assert np.array_equal(formats.read_vector(str(tmp_path / "sub" / "potentials")), v)
    assert np.array_equal(formats.read_vector(p), v)
    pts = rng.random((31, 3))
    formats.write_vector(pts, str(tmp_path / "pts"), "bin")
    assert os.path.getsize(tmp_path / "pts.bin") == 31 * 3 * 8
    assert np.array_equal(formats.read_vector(str(tmp_path / "pts"), dim=3, file_mode="bin"), pts)
    formats.write_vector(pts, str(tmp_path / "pts"))
    line = open(tmp_path / "pts.txt").readline().split()
    assert len(line) == 3
    assert np.array_equal(formats.read_vector(str(tmp_path / "pts.txt"), dim=3), pts)
    with pytest.raises(FileNotFoundError):
        formats.read_vector(str(tmp_path / "missing"))
    with pytest.raises(ValueError):
        formats.write_vector(v, str(tmp_path / "x"), "vtk")


def test_hierarchy_level_sizes():
    # level 1 = min(n, max), then /4 per level down to min, at most 5 levels
    assert ml.hierarchy_level_sizes(100000, 100, 1000) == [100000, 1000, 250]
    assert ml.hierarchy_level_sizes(100000, 10, 100000) == [100000, 100000, 25000, 6250, 1562]
    assert ml.hierarchy_level_sizes(50, 100, 1000) == [50]


def test_hierarchy_files_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    y = rng.random((400, 2))
    nu = rng.dirichlet(np.ones(400))
    d = str(tmp_path / "h")
    pts, dens, par = ml.generate_target_hierarchy(y, nu, d, sizes=[400, 60, 12], seed=2)
    assert [len(p) for p in pts] == [400, 60, 12]
    files = formats.get_target_hierarchy_files(d)
    assert [formats.level_of(f[0]) for f in files] == [2, 1, 0]   # coarsest first
    assert formats.count_hierarchy_levels(d) == 3
    for k in range(3):
        assert np.array_equal(formats.read_vector(files[2 - k][0], dim=2), pts[k])
        assert np.array_equal(formats.read_vector(files[2 - k][1]), dens[k])
    # weights summed per cluster; children lists invert the parent map
    for k in (1, 2):
        assert np.isclose(dens[k].sum(), 1.0, rtol=0, atol=1e-14)
        ch = formats.load_hierarchy_children(d, k - 1)
        assert len(ch) == len(pts[k])
        back = np.full(len(pts[k - 1]), -1)
        for c, lst in enumerate(ch):
            back[lst] = c
        assert np.array_equal(back, par[k - 1])
        assert np.allclose([dens[k - 1][lst].sum() for lst in ch], dens[k], rtol=1e-13, atol=0)
        parents = formats.read_index_lists(os.path.join(d, f"level_{k - 1}_parents.txt"))
        assert [p[0] for p in parents] == par[k - 1].tolist()
    # a k-means fixed point: every point sits with its nearest center
    d2 = np.sum((pts[0][:, None, :] - pts[1][None, :, :]) ** 2, axis=2)
    assert np.array_equal(np.argmin(d2, axis=1), par[0])
    with pytest.raises(FileNotFoundError):
        formats.load_hierarchy_children(d, 2)
    with pytest.raises(RuntimeError):
        formats.get_target_hierarchy_files(str(tmp_path / "none"))


def test_direct_assignment_children_take_parent_value():
    class _Dim:
        dim = 2
    prev = np.array([1.5, -2.0, 0.25])
    children = [[], [[0, 3], [1], [2, 4]]]
    out = ml.assign_potentials_by_hierarchy(_Dim(), 2, 1, prev, np.zeros(12), None, None,
                                            children, False, 0.0, 0.01, True)
    assert out.tolist() == [1.5, -2.0, 0.25, 1.5, 0.25, 0.0]
    same = ml.assign_potentials_by_hierarchy(_Dim(), 1, 1, prev, np.zeros(6), None, None,
                                             children, False, 0.0, 0.01, True)
    assert same.tolist() == prev.tolist()
