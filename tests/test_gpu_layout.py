"""The target layout built on the device (sdot_set_target: stable radix sorts
by row-major bin and by Morton key, bin offsets) against a numpy restatement
of the same rules: the bin of a point is floor((v - o) / h) clamped to the
grid, its Morton key interleaves 21 (3D) or 31 (2D) bits of its position in
the bounding box, and both orders are stable (ties keep caller order).  Exact
equality of every position and offset.  Cases: duplicates, points on bin
faces, one point, sizes that are not multiples of 64, 2D and 3D, a 500,000-
target set.  Then an evaluation against the oracle on a layout with many
ties."""
import ctypes as C
import sys

import numpy as np
import pytest

import semidiscreteot_amd as sd
from semidiscreteot_amd import _lib as L
from oracle import oracle as O
from tests.cases import structured_case

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-layout] {msg}", flush=True)
    sys.stdout.flush()


def device_layout(s, n):
    pos_r = np.zeros(n, dtype=np.int32)
    pos_m = np.zeros(n, dtype=np.int32)
    nb = C.c_int64()
    grid = np.zeros(7)
    L.check(s.lib.sdot_target_layout(s._h, None, None, None, C.byref(nb),
                                     grid.ctypes.data_as(C.POINTER(C.c_double))))
    tstart = np.zeros(nb.value + 1, dtype=np.int32)
    L.check(s.lib.sdot_target_layout(s._h, pos_r.ctypes.data, pos_m.ctypes.data,
                                     tstart.ctypes.data, C.byref(nb),
                                     grid.ctypes.data_as(C.POINTER(C.c_double))))
    return pos_r, pos_m, tstart, grid


def numpy_layout(y, dim, grid):
    o, h, nbin = grid[:3], grid[3], grid[4:7].astype(np.int64)
    inv_h = 1.0 / h
    b = np.zeros((y.shape[0], 3), dtype=np.int64)
    for d in range(dim):
        f = np.floor((y[:, d] - o[d]) * inv_h)
        bd = np.where(f >= 0.0, f, 0.0)
        bd = np.where(bd >= nbin[d], nbin[d] - 1, bd)
        b[:, d] = bd.astype(np.int64)
    bins = (b[:, 2] * nbin[1] + b[:, 1]) * nbin[0] + b[:, 0]
    r2o = np.argsort(bins, kind="stable")
    pos_r = np.empty_like(r2o)
    pos_r[r2o] = np.arange(r2o.size)
    n_bins = int(nbin[0] * nbin[1] * nbin[2])
    tstart = np.searchsorted(bins[r2o], np.arange(n_bins + 1), side="left")
    bits = 21 if dim == 3 else 31
    scale = float(((1 << (bits - 1)) - 1) * 2)
    lo, hi = y.min(0), y.max(0)
    q = []
    for d in range(dim):
        ext = hi[d] - lo[d]
        u = (y[:, d] - lo[d]) / ext if ext > 0 else np.zeros(y.shape[0])
        u = np.where(u > 0.0, u, 0.0)
        u = np.where(u < 1.0, u, 1.0)
        q.append((u * scale).astype(np.uint64))
    key = np.zeros(y.shape[0], dtype=np.uint64)
    one = np.uint64(1)
    for bit in range(bits - 1, -1, -1):
        for d in range(dim - 1, -1, -1):
            key = (key << one) | ((q[d] >> np.uint64(bit)) & one)
    m2o = np.argsort(key, kind="stable")
    pos_m = np.empty_like(m2o)
    pos_m[m2o] = np.arange(m2o.size)
    return pos_r, pos_m, tstart


def cases():
    rng = np.random.default_rng(3)
    out = []
    y = rng.random((1000, 3))
    y[500:600] = y[:100]                       # duplicates: ties in both orders
    out.append(("3d duplicates", 3, y))
    g = np.arange(0, 1.0001, 0.125)            # points on bin faces of a dyadic grid
    y = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    out.append(("3d lattice", 3, y[rng.permutation(len(y))]))
    out.append(("one point", 3, np.array([[0.3, 0.4, 0.5]])))
    out.append(("2d 777", 2, rng.random((777, 2)) * [2.0, 0.5] - 0.25))
    out.append(("3d flat", 3, np.c_[rng.random((300, 2)), np.full(300, 0.5)]))
    y = rng.standard_normal((500_000, 3)) * 0.1 + 0.5
    out.append(("3d 500000", 3, y))
    return out


@pytest.mark.parametrize("k", range(6))
def test_device_layout_matches_numpy(k):
    name, dim, y = cases()[k]
    src, _ = structured_case(dim, 4, 2, 16, seed=1)
    s = sd.SotSolver(dim)
    s.setup_source(src.x, src.w, src.center, src.nq)
    n = y.shape[0]
    s.setup_target(np.ascontiguousarray(y).ravel(), np.full(n, 1.0 / n))
    pos_r, pos_m, tstart, grid = device_layout(s, n)
    r_ref, m_ref, t_ref = numpy_layout(y, dim, grid)
    say(f"{name}: n {n}, bins {tstart.size - 1}, grid {grid[4:7].astype(int).tolist()}")
    assert np.array_equal(pos_r, r_ref)
    assert np.array_equal(pos_m, m_ref)
    assert np.array_equal(tstart, t_ref)
    s.close()


def test_evaluation_on_tied_layout_vs_oracle():
    """Targets on a dyadic lattice (every point on bin faces, repeated
    points): FP64 evaluation and candidate statistics against the oracle."""
    g = np.arange(0.0625, 1.0, 0.125)
    y = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    y = np.concatenate([y, y[:64]])
    n = y.shape[0]
    nu = np.random.default_rng(2).dirichlet(np.full(n, 5.0))
    src, _ = structured_case(3, 8, 2, 16, seed=4, gmm_source=True)
    tgt = O.Target(y.ravel(), nu)
    s = sd.SotSolver(3)
    s.setup_source(src.x, src.w, src.center, src.nq)
    s.setup_target(tgt.y, tgt.nu)
    s.configure(sd.SolverParameters(epsilon=1e-2, use_log_sum_exp_trick=True))
    psi = 0.01 * np.random.default_rng(5).standard_normal(n)
    for R in (0.3, np.finfo(float).max):
        F, gr, _ = s.evaluate_functional_R(R, psi)
        Fo, go, _ = O.evaluate_R(src, tgt, O.params(epsilon=1e-2, use_lse=1), R, psi)
        say(f"R {R:.3g}: F rel {abs(F - Fo) / abs(Fo):.2e}")
        assert abs(F - Fo) <= 1e-10 * abs(Fo)
        assert np.max(np.abs(gr - go)) <= 1e-10 * np.max(np.abs(go) + nu)
        cnt, h, _ = s.candidate_stats(R)
        co, ho = O.candidate_stats(src, tgt, R)
        assert np.array_equal(cnt, co) and np.array_equal(h, ho)
    s.close()
