"""The benchmark configurations at full size on the GPU.

  C2  3D Gaussian mixture, 32^3 hexes QGauss(2) (262,144 q-points), 50,000
      targets, eps 1e-3, at the benchmark's potential (60 untruncated FP64
      L-BFGS iterations from 0) and calibrated radius (discarded mass
      < 1e-12): FP64 against the oracle (|F - F_ref| <= 1e-10 |F_ref|,
      max_j |g_j - g_ref_j| <= 1e-10 max_j (|g_ref_j| + nu_j)), per-cell
      candidate counts and hashes equal to the oracle's, and FP32_LSE, cold
      and seeded, within 1e-6 of max(nu_j, m_j) (F within 1e-8).
  C3  combined multilevel 16^3 -> 32^3 -> 64^3 hexes, 1,000 -> 10,000 ->
      100,000 targets, eps 1e-1 -> 1e-2 -> 1e-3: every level converges and the
      FP64 relative marginal error max_j |g_j| / nu_j at the final potential
      and the last solve's radius is <= 1e-6.
  C4  Delaunay ball (200,000 points, 1.34M tetrahedra, 5.37M q-points),
      500,000 targets, eps 5e-4, at a potential from 30 FP64 iterations and
      its pointwise radius: properties that hold at any size -- FP32_LSE
      within 1e-6 of FP64, sum_j g_j equal to the mass of the q-points with a
      candidate minus 1, candidate counts and hashes of 256 sampled cells equal
      to a numpy brute-force scan of every target, and FP64 on 20,000 sampled
      cells (all 500,000 targets) against the oracle: assembled marginals
      within 1e-10 of their maximum, F within 1e-10.
  C5  three requests of the seeded serving trace (47,790 / 7,613 / 1,174
      targets) on the C2 mesh, served together and one at a time: bitwise
      equal potentials, and each FP64-oracle relative marginal error <= 1e-6."""
import itertools
import sys
import time

import numpy as np
import pytest

import semidiscreteot_amd as sd
from semidiscreteot_amd import workloads
from oracle import oracle as O

pytestmark = pytest.mark.gpu
DMAX = np.finfo(float).max


def say(msg):
    print(f"[gpu-cfg] {msg}", flush=True)
    sys.stdout.flush()


def oracle_source(src):
    return O.Source(src.dim, src.nq, src.x, src.w, np.ones_like(src.w), src.centers)


def solver_for(wl):
    s = sd.SotSolver(wl.source.dim)
    s.setup_source(wl.source.x, wl.source.w, wl.source.centers, wl.source.nq)
    s.setup_target(wl.y, wl.nu)
    return s


def relm(g, go, nu):
    return float(np.max(np.abs(g - go) / np.maximum(nu, go + nu)))


def scan_stats(centers, y, R, cells):
    """Candidate count and hash of each listed cell by a scan of every target
    with the candidate rule sqrt(((y0-c0)^2 + (y1-c1)^2) + (y2-c2)^2) <= R
    (every numpy operation rounded on its own, no contraction); the hash is
    the sum over candidates of mix(j + 1), mix(v) = (v * 0x9E3779B97F4A7C15)
    ^ (that >> 29), modulo 2^64."""
    dim = centers.shape[1]
    cnt = np.zeros(len(cells), dtype=np.int64)
    h = np.zeros(len(cells), dtype=np.uint64)
    k = np.uint64(0x9E3779B97F4A7C15)
    for i, c in enumerate(cells):
        d = y[:, 0] - centers[c, 0]
        s = d * d
        for a in range(1, dim):
            d = y[:, a] - centers[c, a]
            s = s + d * d
        idx = np.nonzero(np.sqrt(s) <= R)[0].astype(np.uint64)
        v = (idx + np.uint64(1)) * k
        v ^= v >> np.uint64(29)
        cnt[i] = idx.size
        h[i] = np.sum(v, dtype=np.uint64)
    return cnt, h


def test_c2_gmm3d_full_size_vs_oracle():
    t0 = time.perf_counter()
    wl = workloads.gmm_3d()
    s = solver_for(wl)
    prm = sd.SolverParameters(epsilon=wl.epsilon, tau=0.0, use_log_sum_exp_trick=True,
                              verbose_output=False, solver_control_type="componentwise",
                              tolerance=1e-6, max_iterations=60, precision="fp64")
    s.configure(prm)
    psi = s.solve()
    lz_full = s.q_log_partition(DMAX, psi)
    lo, hi = 0.02, float(np.sqrt(3.0))
    for _ in range(14):
        mid = 0.5 * (lo + hi)
        if s.discarded_mass(mid, psi, wl.source.w, lz_full) < 1e-12:
            hi = mid
        else:
            lo = mid
    R = hi
    say(f"C2: psi from 60 iterations, R = {R:.6f} ({time.perf_counter() - t0:.1f}s)")
    F, g, C = s.evaluate_functional_R(R, psi)
    cnt, hsh, pairs = s.candidate_stats(R)
    t1 = time.perf_counter()
    src, tgt = oracle_source(wl.source), O.Target(wl.y, wl.nu)
    op = O.params(epsilon=wl.epsilon, use_lse=1)
    Fo, go, Co = O.evaluate_R(src, tgt, op, R, psi)
    cnt_o, hsh_o = O.candidate_stats(src, tgt, R)
    say(f"C2: oracle evaluation and candidate scan {time.perf_counter() - t1:.1f}s; "
        f"{pairs} pairs")
    rF, rg = abs(F - Fo) / abs(Fo), np.max(np.abs(g - go)) / np.max(np.abs(go) + wl.nu)
    say(f"C2 FP64: F rel {rF:.3e}, grad rel {rg:.3e}, C rel {abs(C - Co) / abs(Co):.3e}")
    assert rF <= 1e-10 and rg <= 1e-10 and abs(C - Co) <= 1e-10 * abs(Co)
    assert np.array_equal(cnt, cnt_o) and np.array_equal(hsh, hsh_o)
    assert pairs == int(np.sum(cnt_o)) * wl.source.nq
    # FP32_LSE: a cold evaluation, then one seeded from it at a nearby potential
    prm.precision = "fp32_lse"
    s.configure(prm)
    psi_b = psi + 0.05 * wl.epsilon * np.random.default_rng(7).standard_normal(psi.size)
    F32, g32, _ = s.evaluate_functional_R(R, psi)
    F32b, g32b, _ = s.evaluate_functional_R(R, psi_b)
    Fob, gob, _ = O.evaluate_R(src, tgt, op, R, psi_b)
    e1, e2 = relm(g32, go, wl.nu), relm(g32b, gob, wl.nu)
    say(f"C2 FP32_LSE: cold grad/max(nu,m) {e1:.3e} F rel {abs(F32 - Fo) / abs(Fo):.3e}; "
        f"seeded {e2:.3e} F rel {abs(F32b - Fob) / abs(Fob):.3e}")
    assert e1 <= 1e-6 and e2 <= 1e-6
    assert abs(F32 - Fo) <= 1e-8 * abs(Fo) and abs(F32b - Fob) <= 1e-8 * abs(Fob)
    s.close()


def test_c3_multilevel_end_to_end(tmp_path):
    from semidiscreteot_amd import formats
    t0 = time.perf_counter()
    wl = workloads.multilevel_3d()
    pts, dens, par = wl.hierarchy
    formats.write_target_hierarchy(str(tmp_path), pts, dens, par)
    say(f"C3: workload built in {time.perf_counter() - t0:.1f}s")
    s = sd.SotSolver(3)
    prm = sd.SolverParameters(epsilon=wl.epsilon, tau=wl.tau, tolerance=1e-6,
                              solver_control_type="componentwise", max_iterations=3000,
                              use_log_sum_exp_trick=True, precision="fp32_lse",
                              line_search="noise_aware", preconditioner="diagonal",
                              use_epsilon_scaling=True, epsilon_scaling_factor=wl.epsilon_factor,
                              epsilon_scaling_steps=wl.epsilon_steps, quadrature_order=2,
                              verbose_output=False)
    ml = sd.MultilevelParameters(target_hierarchy_dir=str(tmp_path))
    t1 = time.perf_counter()
    psi, recs = sd.run_combined_multilevel(s, wl.sources, prm, ml)
    for r in recs:
        say(f"C3 level {r.level}: {r.n_cells} cells, {r.n_targets} targets, eps {r.epsilons}, "
            f"{r.status} after {r.iterations} iterations, R {r.threshold:.4f}")
    assert all(r.status == "converged" for r in recs)
    assert [r.n_targets for r in recs] == [1000, 10000, 100000]
    s.setup_target(pts[0].ravel(), dens[0])
    s.configure(sd.SolverParameters(epsilon=wl.epsilon, tau=0.0, use_log_sum_exp_trick=True,
                                    verbose_output=False))
    _, g, _ = s.evaluate_functional_R(recs[-1].threshold, psi)
    err = float(np.max(np.abs(g) / dens[0]))
    say(f"C3: {time.perf_counter() - t1:.1f}s; final FP64 relative marginal error {err:.3e}")
    assert err <= 1e-6
    s.close()


def test_c4_ball_tets_full_size_properties():
    t0 = time.perf_counter()
    wl = workloads.ball_tets()
    n_q = wl.source.x.size // 3
    s = solver_for(wl)
    say(f"C4: {wl.source.n_cells} cells, {n_q} q, {wl.nu.size} targets "
        f"({time.perf_counter() - t0:.1f}s)")
    base = dict(epsilon=wl.epsilon, tau=wl.tau, use_log_sum_exp_trick=True, verbose_output=False,
                solver_control_type="componentwise", tolerance=1e-6, max_iterations=30,
                line_search="noise_aware", preconditioner="diagonal")
    s.configure(sd.SolverParameters(precision="fp64", **base))
    psi = s.solve()
    R = s.distance_threshold(psi)
    F, g, C = s.evaluate_functional_R(R, psi)
    s.configure(sd.SolverParameters(precision="fp32_lse", **base))
    F32, g32, C32 = s.evaluate_functional_R(R, psi)
    e = relm(g32, g, wl.nu)
    say(f"C4: R {R:.4f}; FP32 vs FP64 grad/max(nu,m) {e:.3e}, F rel {abs(F32 - F) / abs(F):.3e}")
    assert e <= 1e-6 and abs(F32 - F) <= 1e-8 * abs(F) and abs(C32 - C) <= 1e-8 * abs(C)
    # mass balance: sum_j m_j is the mass of the q-points with a candidate
    lz = s.q_log_partition(R, psi)
    kept = float(np.sum(wl.source.w[np.isfinite(lz)]))
    bal = abs(float(np.sum(g)) - (kept - float(np.sum(wl.nu))))
    say(f"C4: sum g {np.sum(g):.3e}, kept mass - 1 {kept - np.sum(wl.nu):.3e}, |diff| {bal:.2e}")
    assert bal <= 1e-12
    cnt, hsh, _ = s.candidate_stats(R)
    cells = np.random.default_rng(11).choice(wl.source.n_cells, 256, replace=False)
    c_ref, h_ref = scan_stats(wl.source.centers.reshape(-1, 3), wl.y.reshape(-1, 3), R, cells)
    say(f"C4: sampled cells' candidates {int(c_ref.min())}..{int(c_ref.max())}")
    assert np.array_equal(cnt[cells], c_ref) and np.array_equal(hsh[cells], h_ref)
    s.close()
    # FP64 against the oracle on 20,000 sampled cells with all 500,000 targets
    # (same potential and radius): the assembled marginals m = g + nu
    sub = np.sort(np.random.default_rng(12).choice(wl.source.n_cells, 20000, replace=False))
    nq = wl.source.nq
    qi = (sub[:, None] * nq + np.arange(nq)).ravel()
    xs = wl.source.x.reshape(-1, 3)[qi].ravel()
    ws = wl.source.w[qi]
    cs = wl.source.centers.reshape(-1, 3)[sub].ravel()
    s2 = sd.SotSolver(3)
    s2.setup_source(xs, ws, cs, nq)
    s2.setup_target(wl.y, wl.nu)
    s2.configure(sd.SolverParameters(precision="fp64", **base))
    F2, g2, _ = s2.evaluate_functional_R(R, psi)
    s2.close()
    t1 = time.perf_counter()
    Fo, go, _ = O.evaluate_R(O.Source(3, nq, xs, ws, np.ones_like(ws), cs), O.Target(wl.y, wl.nu),
                             O.params(epsilon=wl.epsilon, use_lse=1), R, psi)
    mo = go + wl.nu
    em = float(np.max(np.abs(g2 - go)) / np.max(mo))
    say(f"C4 sample of 2000 cells: oracle {time.perf_counter() - t1:.1f}s, max |m - m_ref| / "
        f"max m_ref {em:.2e}, F rel {abs(F2 - Fo) / abs(Fo):.2e}")
    assert em <= 1e-10 and abs(F2 - Fo) <= 1e-10 * abs(Fo)


def test_c5_trace_requests_mix_vs_alone_and_oracle():
    wl = workloads.gmm_3d()
    trace = list(itertools.islice(workloads.serving_trace(), 5))
    reqs = [trace[i] for i in (0, 2, 4)]
    assert reqs[0].nu.size > 45000
    s = sd.SotSolver(3)
    s.setup_source(wl.source.x, wl.source.w, wl.source.centers, wl.source.nq)
    kw = dict(control="componentwise", max_iterations=1000, preconditioner="diagonal")

    def serve(rs, stagger):
        with sd.SotServer(s, precision="fp32_lse", max_slots=8, hbm_budget_bytes=8 << 30) as srv:
            ids = []
            for k, r in enumerate(rs):
                ids.append(srv.submit(r.y, r.nu, r.epsilon, r.tolerance, **kw))
                if stagger and k == 0:
                    srv.step()
                    srv.step()
            out = srv.run(ids)
            return [out[i] for i in ids]

    t0 = time.perf_counter()
    mix = serve(reqs, True)
    alone = [serve([r], False)[0] for r in reqs]
    say(f"C5: served in {time.perf_counter() - t0:.1f}s")
    src = oracle_source(wl.source)
    for r, (rm, pm), (ra, pa) in zip(reqs, mix, alone):
        t1 = time.perf_counter()
        _, go, _ = O.evaluate_R(src, O.Target(r.y, r.nu), O.params(epsilon=r.epsilon, use_lse=1),
                                DMAX, pm)
        err = float(np.max(np.abs(go) / r.nu))
        say(f"C5 request {r.index} (N {r.nu.size}, eps {r.epsilon}): {rm.status} after "
            f"{rm.iterations} iterations, oracle relerr {err:.3e} "
            f"({time.perf_counter() - t1:.1f}s), bitwise {np.array_equal(pm, pa)}")
        assert rm.status == ra.status == "converged"
        assert np.array_equal(pm, pa) and rm.iterations == ra.iterations
        assert err <= 1e-6
    s.close()
