"""Pins of the FP64 CPU oracle: golden 50-digit vectors, closed forms,
invariants and textbook problems, each chosen so that a plausible slip in the
restatement (dropped linear term, wrong sign or index, transposed operands,
missing max shift) fails at least one of them."""
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import structured_case, tet_case

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "tiny_cases.json")))


def golden_source(c):
    return (O.Source(c["dim"], c["nq"], c["x"], c["rho"], c["jxw"], c["center"]),
            O.Target(c["y"], c["nu"]))


@pytest.mark.parametrize("case", GOLD, ids=[c["name"] for c in GOLD])
def test_oracle_matches_50_digit_golden(case):
    src, tgt = golden_source(case)
    prm = O.params(epsilon=case["eps"], use_lse=case["use_lse"], n_threads=2)
    F, g, C = O.evaluate_R(src, tgt, prm, case["R"], np.array(case["psi"]))
    Fg = float(case["F"])
    gg = np.array([float(v) for v in case["g"]])
    assert abs(F - Fg) <= 1e-13 * max(1.0, abs(Fg))
    scale = np.max(np.abs(gg + np.array(case["nu"])))
    assert np.max(np.abs(g - gg)) <= 1e-13 * scale
    assert abs(C - float(case["C"])) <= 1e-12 * abs(float(case["C"]))
    cnt, h = O.candidate_stats(src, tgt, case["R"])
    assert cnt.tolist() == case["count"]
    assert [str(v) for v in h.tolist()] == case["hash"]


def test_gauss_legendre_textbook():
    p, w = O.gauss_legendre01(2)
    assert np.allclose(p, [0.5 - 0.5 / np.sqrt(3), 0.5 + 0.5 / np.sqrt(3)], atol=1e-15)
    assert np.allclose(w, [0.5, 0.5], atol=1e-15)
    p, w = O.gauss_legendre01(3)
    assert np.allclose(p, [0.5 - 0.5 * np.sqrt(0.6), 0.5, 0.5 + 0.5 * np.sqrt(0.6)], atol=1e-15)
    assert np.allclose(w, [5 / 18, 8 / 18, 5 / 18], atol=1e-15)
    for n in range(1, 8):  # exact for degree 2n-1
        p, w = O.gauss_legendre01(n)
        assert abs(np.sum(w * p ** (2 * n - 1)) - 1.0 / (2 * n)) < 1e-14


def test_structured_source_integrates_polynomials():
    src = O.structured_source(3, [3, 4, 5], [0.0, -1.0, 2.0], [1.0, 1.0, 3.0], 2, normalize=False)
    x = src.x.reshape(-1, 3)
    assert abs(src.jxw.sum() - 2.0) < 1e-13
    # int x0 y1^2 z2 over the box = (1/2) * (2/3) * (5/2)
    assert abs(np.sum(src.jxw * x[:, 0] * x[:, 1] ** 2 * x[:, 2]) - 0.5 * (2 / 3) * 2.5) < 1e-13
    # bilinear nodal density is reproduced exactly at quadrature points
    g = [np.linspace(0, 1, 4), np.linspace(-1, 1, 5), np.linspace(2, 3, 6)]
    Z, Y, X = np.meshgrid(g[2], g[1], g[0], indexing="ij")
    f = lambda a, b, c: 1 + a + 2 * b + 3 * c + a * b * c
    s2 = O.structured_source(3, [3, 4, 5], [0.0, -1.0, 2.0], [1.0, 1.0, 3.0], 2,
                             f(X, Y, Z).ravel(), normalize=False)
    assert np.allclose(s2.rho, f(x[:, 0], x[:, 1], x[:, 2]), atol=1e-13)
    c = src.center.reshape(-1, 3)
    assert np.allclose(c[0], [1 / 6, -0.75, 2.1])


def test_simplex_source_volume_and_rule():
    verts = np.array([[0, 0, 0], [2, 0, 0], [0, 3, 0], [0, 0, 1.5]], dtype=float)
    src = O.simplex_source(3, verts.ravel(), np.array([0, 1, 2, 3]), 2, normalize=False)
    assert abs(src.jxw.sum() - 2 * 3 * 1.5 / 6) < 1e-14
    x = src.x.reshape(-1, 3)
    # the 4-point rule is exact for quadratics:
    # int_T x0^2 = V (sum_i v_i^2 + sum_{i<j} v_i v_j) / 10 with v = x0 of the vertices
    vol = 1.5
    exact_xx = vol * (2.0 ** 2) / 10.0
    assert abs(np.sum(src.jxw * x[:, 0] ** 2) - exact_xx) < 1e-13


def test_single_target_closed_form():
    """N=1, nu=1: g = mass - 1 = 0 and F = -1/2 sum_q w_q |x_q - y|^2, which for a
    uniform unit square is -(|c-y|^2 + 1/6)/2 (QGauss(3) integrates it exactly)."""
    src = O.structured_source(2, [8, 8], [0, 0], [1, 1], 3)
    for y, psi in [((0.3, 0.6), 0.0), ((1.4, -0.2), 0.7)]:
        tgt = O.Target(np.array(y), [1.0])
        for lse in (0, 1):
            F, g, _ = O.evaluate_R(src, tgt, O.params(epsilon=1e-2, use_lse=lse), 1e300, [psi])
            cf = -0.5 * ((0.5 - y[0]) ** 2 + (0.5 - y[1]) ** 2 + 1 / 6)
            assert abs(F - cf) < 1e-13
            assert abs(g[0]) < 1e-13


def test_quadrant_symmetry_zero_gradient():
    src = O.structured_source(2, [8, 8], [0, 0], [1, 1], 3)
    y = np.array([[0.25, 0.25], [0.75, 0.25], [0.25, 0.75], [0.75, 0.75]])
    tgt = O.Target(y.ravel(), np.full(4, 0.25))
    for eps in (1e-1, 1e-2, 1e-3):
        F, g, _ = O.evaluate_R(src, tgt, O.params(epsilon=eps, use_lse=1), 1e300, np.zeros(4))
        assert np.max(np.abs(g)) < 1e-13


def test_invariants_and_finite_differences():
    src, tgt = structured_case(2, 6, 3, 15, seed=11, gmm_source=True)
    rng = np.random.default_rng(5)
    psi = 0.02 * rng.standard_normal(tgt.n)
    prm = O.params(epsilon=2e-2, use_lse=1)
    F, g, _ = O.evaluate_R(src, tgt, prm, 1e300, psi)
    assert abs(g.sum()) < 1e-13                     # sum g = mass - sum nu = 0
    F2, g2, _ = O.evaluate_R(src, tgt, prm, 1e300, psi + 0.3)
    assert abs(F2 - F) < 1e-13 and np.max(np.abs(g2 - g)) < 1e-13  # shift invariance
    h = 1e-6
    fd = [(O.evaluate_R(src, tgt, prm, 1e300, psi + h * e)[0] -
           O.evaluate_R(src, tgt, prm, 1e300, psi - h * e)[0]) / (2 * h) for e in np.eye(tgt.n)]
    assert np.max(np.abs(np.array(fd) - g)) < 1e-8
    # plain and LSE modes agree where nothing underflows
    Fp, gp, Cp = O.evaluate_R(src, tgt, O.params(epsilon=2e-2, use_lse=0), 1e300, psi)
    assert abs(Fp - F) < 1e-13 and np.max(np.abs(gp - g)) < 1e-14


def test_hessian_fd_symmetry_kernel():
    src, tgt = structured_case(2, 5, 2, 12, seed=3)
    psi = 0.01 * np.random.default_rng(1).standard_normal(tgt.n)
    prm = O.params(epsilon=3e-2, use_lse=1)
    H = O.hessian(src, tgt, prm, 1e300, psi)
    h = 1e-6
    fd = np.array([(O.evaluate_R(src, tgt, prm, 1e300, psi + h * e)[1] -
                    O.evaluate_R(src, tgt, prm, 1e300, psi - h * e)[1]) / (2 * h)
                   for e in np.eye(tgt.n)])
    assert np.max(np.abs(fd - H)) < 1e-7
    assert np.max(np.abs(H - H.T)) < 1e-15
    assert np.max(np.abs(H.sum(1))) < 1e-12
    assert np.min(np.linalg.eigvalsh(H)) > -1e-12


def test_thresholds_closed_form():
    src, tgt = structured_case(2, 4, 2, 10, seed=2, dirichlet=5)
    psi = np.linspace(-0.1, 0.2, tgt.n)
    st = O.new_state()
    eps, tau = 1e-2, 1e-8
    R = O.distance_threshold(src, tgt, O.params(epsilon=eps, tau=tau), st, psi)
    assert abs(R - np.sqrt(-2 * eps * np.log(tau / tgt.nu.min()) + 2 * 0.2)) < 1e-15
    st.C_global, st.global_functional, st.covering_radius = 3.0, -0.05, 0.01
    R = O.distance_threshold(src, tgt, O.params(epsilon=eps, tau=tau, threshold="integral"), st, psi)
    assert abs(R - np.sqrt(2 * 0.2 + 2 * eps * np.log(eps * 3.0 / (tau * 0.05)))) < 1e-15
    st.covering_radius, st.min_target_density = 0.3, 0.0
    R = O.distance_threshold(src, tgt, O.params(epsilon=eps, tau=tau, threshold="geometric"), st, psi)
    r2 = 0.09 + 2 * 0.3 + 2 * eps * np.log(eps / (tgt.nu.min() * tau * 0.05))
    assert abs(R - np.sqrt(r2)) < 1e-15
    # tau <= 0 disables truncation
    assert O.distance_threshold(src, tgt, O.params(tau=0.0), O.new_state(), psi) == np.finfo(float).max


def test_candidate_sets_brute_force():
    src, tgt = structured_case(3, 4, 2, 60, seed=8, gmm_targets=True)
    R = 0.31
    cnt, h = O.candidate_stats(src, tgt, R)
    c = src.center.reshape(-1, 3)
    y = tgt.y.reshape(-1, 3)
    d = np.sqrt(((y[None, :, :] - c[:, None, :]) ** 2).sum(-1))
    assert cnt.tolist() == (d <= R).sum(1).tolist()


def test_bfgs_textbook_problems():
    A = np.array([[4.0, 1.0, 0.5], [1.0, 3.0, 0.2], [0.5, 0.2, 2.0]])
    b = np.array([1.0, -2.0, 0.5])
    x, info = O.bfgs_generic(lambda x: (0.5 * x @ A @ x - b @ x, A @ x - b), np.zeros(3), tol=1e-12)
    assert info["status"] == "converged"
    assert np.allclose(x, np.linalg.solve(A, b), atol=1e-11)
    rb = lambda x: ((1 - x[0]) ** 2 + 100 * (x[1] - x[0] ** 2) ** 2,
                    np.array([-2 * (1 - x[0]) - 400 * x[0] * (x[1] - x[0] ** 2),
                              200 * (x[1] - x[0] ** 2)]))
    x, info = O.bfgs_generic(rb, np.array([-1.2, 1.0]), max_iterations=2000, tol=1e-10)
    assert info["status"] == "converged" and np.allclose(x, [1, 1], atol=1e-8)


def test_solve_reaches_componentwise_tolerance():
    src, tgt = structured_case(2, 8, 2, 30, seed=4, gmm_source=True, dirichlet=5)
    prm = O.params(epsilon=2e-2, tau=1e-12, use_lse=1, control="componentwise",
                   tolerance=1e-6, max_iterations=500)
    psi, info = O.solve(src, tgt, prm)
    assert info["status"] == "converged"
    F, g, _ = O.evaluate_R(src, tgt, prm, 1e300, psi)
    assert np.max(np.abs(g) / tgt.nu) <= 1e-6 * 1.0001
    prm1 = O.params(epsilon=2e-2, use_lse=1, max_iterations=1, tolerance=1e-12)
    _, info1 = O.solve(src, tgt, prm1)
    assert info1["status"] == "max_iterations" and info1["iterations"] == 1


def test_tet_source_mass():
    src, tgt, _ = tet_case(300, 20, seed=3)
    assert abs(src.w.sum() - 1.0) < 1e-12
    assert src.nq == 4


# ---- softmax potential transfer (SoftmaxRefinement.cc:61-280) ----

REFINE = json.load(open(os.path.join(HERE, "golden", "refine_cases.json")))


@pytest.mark.parametrize("case", REFINE, ids=[c["name"] for c in REFINE])
def test_refinement_matches_50_digit_golden(case):
    """Truncated and untruncated transfers with a child listed under two
    parents, an out-of-range child index, and (tight case) fine points that
    receive nothing and keep 0."""
    src = O.Source(case["dim"], case["nq"], case["x"], case["rho"], case["jxw"], case["center"])
    coarse = O.Target(case["y_coarse"], case["nu_coarse"])
    psi = O.softmax_refinement(src, case["y_fine"], coarse, case["psi_coarse"], case["children"],
                               case["lambda"], case["R"], case["use_lse"], n_threads=2)
    ref = np.array([float(v) for v in case["psi_fine"]])
    assert np.array_equal(psi == 0.0, ref == 0.0)
    assert np.max(np.abs(psi - ref)) <= 1e-15


@pytest.mark.parametrize("lse", [True, False])
def test_refinement_identity_hierarchy(lse):
    """Fine level = coarse level, each point its own child, no truncation:
    sum_q w_q exp(-d^2/2l)/Z_q = (g_c + nu_c) / nu_c * exp(-psi_c/l), so
    psi_f = psi_c - l log(1 + g_c / nu_c) with g the gradient the evaluation
    assembles at the same potential and epsilon = l."""
    src, tgt = structured_case(2, 10, 2, 30, 7, gmm_source=True, dirichlet=3.0)
    rng = np.random.default_rng(8)
    psi = 0.02 * rng.standard_normal(tgt.n)
    lam = 0.015
    prm = O.params(epsilon=lam, use_lse=lse, n_threads=2)
    _, g, _ = O.evaluate_R(src, tgt, prm, np.finfo(float).max, psi)
    children = [[i] for i in range(tgt.n)]
    pf = O.softmax_refinement(src, tgt.y, tgt, psi, children, lam, np.finfo(float).max, lse)
    assert np.max(np.abs(pf - (psi - lam * np.log1p(g / tgt.nu)))) <= 1e-14
    # a child listed under a second parent receives its sum twice: -l log 2 more
    children[3].append(5)
    pf2 = O.softmax_refinement(src, tgt.y, tgt, psi, children, lam, np.finfo(float).max, lse)
    assert abs(pf2[5] - (pf[5] - lam * np.log(2.0))) <= 1e-14
    assert np.array_equal(np.delete(pf2, 5), np.delete(pf, 5))


def test_refinement_threshold_below_every_distance():
    """No coarse point within the threshold of any cell center: nothing is
    transferred and every fine potential stays 0."""
    src, tgt = structured_case(2, 6, 2, 9, 3)
    y = np.full(tgt.n * 2, 5.0)
    far = O.Target(y, tgt.nu)
    pf = O.softmax_refinement(src, tgt.y, far, np.zeros(tgt.n), [[i] for i in range(tgt.n)],
                              0.01, 1.0, True)
    assert np.all(pf == 0.0)


# ---- weighted barycenters of the Lloyd step (SotSolver.cc:871-971) ----

BARY = json.load(open(os.path.join(HERE, "golden", "bary_cases.json")))


@pytest.mark.parametrize("case", BARY, ids=[c["name"] for c in BARY])
def test_barycenters_match_50_digit_golden(case):
    src = O.Source(case["dim"], case["nq"], case["x"], case["rho"], case["jxw"], case["center"])
    tgt = O.Target(case["y"], case["nu"])
    b = O.barycenters_R(src, tgt, O.params(epsilon=case["eps"], use_lse=True, n_threads=2),
                        case["R"], case["psi"])
    ref = np.array([[float(v) for v in r] for r in case["bary"]])
    assert np.max(np.abs(b - ref)) <= 1e-14 * max(1.0, np.max(np.abs(ref)))


def test_barycenters_closed_forms():
    """One target of weight 1 takes the whole source: its barycenter is
    sum_q w_q x_q.  Moving source and targets by v moves b_j by
    v (1 + g_j / nu_j), the plan's column mass over nu_j, with g the gradient
    the evaluation assembles (no truncation)."""
    src, _ = structured_case(2, 8, 2, 1, 3, gmm_source=True)
    one = O.Target(np.array([0.3, 0.6]), np.array([1.0]))
    b = O.barycenters_R(src, one, O.params(epsilon=0.01, use_lse=True), np.finfo(float).max,
                        np.zeros(1))
    xs = src.x.reshape(-1, 2)
    assert np.max(np.abs(b[0] - (src.w[:, None] * xs).sum(0))) <= 1e-14
    src3, tgt3 = structured_case(3, 6, 2, 40, 9, gmm_source=True, dirichlet=4.0)
    psi = 0.01 * np.random.default_rng(9).standard_normal(tgt3.n)
    prm = O.params(epsilon=0.02, use_lse=True, n_threads=2)
    dmax = np.finfo(float).max
    b0 = O.barycenters_R(src3, tgt3, prm, dmax, psi)
    _, g, _ = O.evaluate_R(src3, tgt3, prm, dmax, psi)
    v = np.array([0.25, -0.5, 0.125])
    moved = O.Source(3, src3.nq, (src3.x.reshape(-1, 3) + v).ravel(), src3.rho, src3.jxw,
                     (src3.center.reshape(-1, 3) + v).ravel())
    b1 = O.barycenters_R(moved, O.Target((tgt3.y.reshape(-1, 3) + v).ravel(), tgt3.nu), prm,
                         dmax, psi)
    pred = b0 + v[None, :] * (1.0 + g / tgt3.nu)[:, None]
    assert np.max(np.abs(b1 - pred)) <= 1e-12


# ---- conditional densities and transport maps ----

POINTS = json.load(open(os.path.join(HERE, "golden", "points_cases.json")))


@pytest.mark.parametrize("case", POINTS, ids=[c["name"] for c in POINTS])
def test_point_queries_match_50_digit_golden(case):
    """Conditional densities at points with threshold R (zero where the
    requested target is outside), barycentric map with a strict truncation
    radius and the nearest-target fallback."""
    tgt = O.Target(case["y"], case["nu"])
    cond = O.conditional_density(tgt, case["dim"], case["psi"], case["eps"], case["R"],
                                 case["points"], case["rho"], case["indices"])
    ref = np.array([[float(v) for v in r] for r in case["cond"]])
    assert np.array_equal(cond == 0.0, ref == 0.0)
    assert np.max(np.abs(cond - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))
    b = O.transport_map(1, tgt, case["dim"], case["psi"], case["eps"], case["radius"],
                        case["points"])
    rb = np.array([[float(v) for v in r] for r in case["bary_map"]])
    assert np.max(np.abs(b - rb)) <= 1e-14


def test_point_queries_closed_forms():
    """sum_i nu_i c_i(x) = rho(x) over all targets (no truncation); the modal
    map with psi = -eps log nu is the nearest-target (Voronoi) map; the
    barycentric map of one target is that target, and tends to the modal map
    as eps -> 0 away from ties."""
    rng = np.random.default_rng(0)
    y = rng.random((30, 2))
    nu = rng.dirichlet(np.ones(30))
    psi = 0.01 * rng.standard_normal(30)
    tgt = O.Target(y.ravel(), nu)
    pts = rng.random((200, 2))
    rho = 1.0 + rng.random(200)
    c = O.conditional_density(tgt, 2, psi, 0.02, np.finfo(float).max, pts, rho, np.arange(30))
    assert np.max(np.abs((nu[:, None] * c).sum(0) - rho)) <= 1e-14
    eps = 0.02
    m = O.transport_map(0, tgt, 2, -eps * np.log(nu), eps, 0.0, pts)
    vor = y[np.argmin(((pts[:, None, :] - y[None]) ** 2).sum(2), 1)]
    assert np.array_equal(m, vor)
    one = O.transport_map(1, O.Target(y[:1].ravel(), np.ones(1)), 2, psi[:1], eps, 0.0, pts)
    assert np.array_equal(one, np.repeat(y[:1], len(pts), 0))
    e = 1e-4
    b = O.transport_map(1, tgt, 2, psi, e, 0.0, pts)
    mo = O.transport_map(0, tgt, 2, psi, e, 0.0, pts)
    sc = psi[None, :] - 0.5 * ((pts[:, None, :] - y[None]) ** 2).sum(2) + e * np.log(nu)[None, :]
    srt = np.sort(sc, 1)
    ok = srt[:, -1] - srt[:, -2] > 50 * e
    assert ok.sum() > 50 and np.max(np.abs(b[ok] - mo[ok])) <= 1e-12


# ---- covering radius (SotSolver.cc:431-473) and discarded mass ----

def test_covering_radius_brute_force_and_closed_form():
    # brute force: max over cells of the distance from the center to the
    # nearest target, the reference's sqrt of the running sum of squares
    for dim, n, N, seed in ((2, 7, 40, 1), (3, 5, 60, 2)):
        src, tgt = structured_case(dim, n, 2, N, seed=seed, gmm_targets=True)
        c = src.center.reshape(-1, dim)
        y = tgt.y.reshape(-1, dim)
        d = np.sqrt(((c[:, None, :] - y[None, :, :]) ** 2).sum(-1))
        assert O.covering_radius(src, tgt) == pytest.approx(d.min(1).max(), rel=1e-15, abs=0)
    # closed form: one target at the center of [0,1]^3 cut into n^3 cells: the
    # farthest cell center is a corner cell's, sqrt(3) (1/2 - 1/(2n)) away
    n = 6
    src = O.structured_source(3, [n] * 3, [0.0] * 3, [1.0] * 3, 1)
    tgt = O.Target(np.array([0.5, 0.5, 0.5]), np.array([1.0]))
    assert O.covering_radius(src, tgt) == pytest.approx(np.sqrt(3.0) * (0.5 - 0.5 / n), rel=1e-15)
    # targets at every cell center: every cell is covered at distance 0
    tgt = O.Target(src.center.copy(), np.full(n ** 3, 1.0 / n ** 3))
    assert O.covering_radius(src, tgt) == 0.0


def test_discarded_fraction_brute_force_and_limits():
    src, tgt = structured_case(2, 6, 2, 25, seed=11, gmm_targets=True, dirichlet=3)
    eps, R = 0.02, 0.3
    psi = 0.01 * np.random.default_rng(3).standard_normal(tgt.n)
    q = np.arange(0, src.n_cells * src.nq, 7, dtype=np.int64)
    got = O.discarded_fraction(src, tgt, O.params(epsilon=eps, use_lse=1), R, psi, q)
    x = src.x.reshape(-1, 2)
    c = src.center.reshape(-1, 2)
    y = tgt.y.reshape(-1, 2)
    for k, qi in enumerate(q):
        e = (psi - 0.5 * ((x[qi] - y) ** 2).sum(1)) / eps
        w = tgt.nu * np.exp(e - e.max())
        out = np.sqrt(((y - c[qi // src.nq]) ** 2).sum(1)) > R
        assert got[k] == pytest.approx(w[out].sum() / w.sum(), rel=1e-12, abs=1e-300)
    # no truncation keeps everything; a negative radius drops everything
    assert np.all(O.discarded_fraction(src, tgt, O.params(epsilon=eps, use_lse=1), 10.0, psi, q) == 0)
    assert np.all(O.discarded_fraction(src, tgt, O.params(epsilon=eps, use_lse=1), -1.0, psi, q) == 1)
    # two targets symmetric about a point, equal weights and potentials, a
    # radius that keeps only the nearer one of a q-point on the axis: the
    # fraction is the farther one's share, 1 / (1 + exp(d / eps)) with
    # d = (|x - y_far|^2 - |x - y_near|^2) / 2
    src1 = O.structured_source(2, [1, 1], [0.0, 0.0], [1.0, 1.0], 1)
    tgt1 = O.Target(np.array([0.2, 0.5, 0.9, 0.5]), np.array([0.5, 0.5]))
    d = 0.5 * ((0.5 - 0.9) ** 2 - (0.5 - 0.2) ** 2)
    got1 = O.discarded_fraction(src1, tgt1, O.params(epsilon=0.05, use_lse=1), 0.35, np.zeros(2),
                                np.array([0], dtype=np.int64))
    assert got1[0] == pytest.approx(1.0 / (1.0 + np.exp(d / 0.05)), rel=1e-14)
