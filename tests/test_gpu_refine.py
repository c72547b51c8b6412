"""GPU parity of the softmax potential transfer between target levels
(sdot_softmax_refinement; SoftmaxRefinement::compute_refinement,
SoftmaxRefinement.cc:61-280) against the 50-digit golden values and the FP64
oracle, and of the multilevel driver built on it.

Tolerance (FP64 throughout): |psi_f - psi_ref| <= 1e-11 lambda + 1e-13 |psi_ref|,
i.e. 1e-11 relative on the transferred mass pot_f = exp(-psi_f / lambda); the
set of fine points that receive nothing (psi_f = 0) is identical."""
import json
import os
import sys

import numpy as np
import pytest

import semidiscreteot_amd as sd
from oracle import oracle as O
from tests.cases import gmm_sample, structured_case

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
REFINE = json.load(open(os.path.join(HERE, "golden", "refine_cases.json")))
DMAX = np.finfo(float).max


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def solver_for(src):
    s = sd.SotSolver(src.dim)
    s.setup_source(src.x, src.w, src.center, src.nq)
    return s


def refine_gpu(s, y_fine, coarse, psi_c, children, lam, R, lse):
    r = sd.SoftmaxRefinement(s, R, lse)
    return r.compute_refinement(y_fine, None, coarse.y, coarse.nu, psi_c, lam, 0, [children])


def assert_close(pg, po, lam):
    assert np.array_equal(pg == 0.0, po == 0.0)
    err = np.abs(pg - po)
    tol = 1e-11 * lam + 1e-13 * np.abs(po)
    assert np.all(err <= tol), (float(np.max(err)), lam)


def nearest_parent_children(yc, yf):
    d = np.sum((yf[:, None, :] - yc[None, :, :]) ** 2, axis=2)
    par = np.argmin(d, axis=1)
    return [[int(i) for i in np.flatnonzero(par == c)] for c in range(len(yc))]


@pytest.mark.parametrize("case", REFINE, ids=[c["name"] for c in REFINE])
def test_refine_golden(case):
    say(f"refine golden {case['name']}")
    src = O.Source(case["dim"], case["nq"], case["x"], case["rho"], case["jxw"], case["center"])
    coarse = O.Target(case["y_coarse"], case["nu_coarse"])
    s = solver_for(src)
    pg = refine_gpu(s, np.array(case["y_fine"]), coarse, np.array(case["psi_coarse"]),
                    case["children"], case["lambda"], case["R"], case["use_lse"])
    ref = np.array([float(v) for v in case["psi_fine"]])
    assert_close(pg, ref, case["lambda"])


CASES = [
    # dim, cells per side, qorder, n_coarse, n_fine, lambda, threshold, lse, seed
    (2, 24, 3, 150, 1200, 0.01, 0.3, True, 21),
    (2, 16, 2, 80, 600, 0.05, DMAX, False, 22),
    (3, 10, 2, 200, 1500, 0.02, 0.35, True, 23),
    (3, 8, 2, 60, 500, 0.03, DMAX, True, 24),
    (2, 20, 2, 120, 900, 0.004, 0.08, True, 25),
]


@pytest.mark.parametrize("dim,n,qo,nc,nf,lam,R,lse,seed", CASES)
def test_refine_vs_oracle(dim, n, qo, nc, nf, lam, R, lse, seed):
    say(f"refine vs oracle dim={dim} n={n} nc={nc} nf={nf} lam={lam} R={R:.3g} lse={lse}")
    src, _ = structured_case(dim, n, qo, 4, seed, gmm_source=True)
    rng = np.random.default_rng(seed)
    yf = gmm_sample(rng, nf, dim, 5)
    yc = yf[rng.choice(nf, nc, replace=False)] + 0.01 * rng.standard_normal((nc, dim))
    children = nearest_parent_children(yc, yf)
    nuc = rng.dirichlet(np.full(nc, 3.0))
    coarse = O.Target(yc.ravel(), nuc)
    psi_c = 0.05 * rng.standard_normal(nc)
    po = O.softmax_refinement(src, yf.ravel(), coarse, psi_c, children, lam, R, lse)
    pg = refine_gpu(solver_for(src), yf.ravel(), coarse, psi_c, children, lam, R, lse)
    assert_close(pg, po, lam)


def test_refine_irregular_child_lists():
    """Duplicate children, out-of-range and negative child indices, empty
    lists and fewer lists than coarse points, against the oracle."""
    say("refine irregular child lists")
    src, _ = structured_case(2, 12, 2, 4, 31, gmm_source=True)
    rng = np.random.default_rng(31)
    nc, nf = 40, 300
    yf = rng.random((nf, 2))
    yc = rng.random((nc, 2))
    children = nearest_parent_children(yc, yf)
    children[0] += [5, 5, nf + 7]
    children[1] = []
    children[2].append(-3)
    children = children[:35]  # coarse points 35..39 have no list
    coarse = O.Target(yc.ravel(), rng.dirichlet(np.ones(nc)))
    psi_c = 0.03 * rng.standard_normal(nc)
    for lse in (True, False):
        valid = [[f for f in ch if 0 <= f < nf] for ch in children]
        po = O.softmax_refinement(src, yf.ravel(), coarse, psi_c, valid, 0.02, 0.25, lse)
        pg = refine_gpu(solver_for(src), yf.ravel(), coarse, psi_c, children, 0.02, 0.25, lse)
        assert_close(pg, po, 0.02)


def test_refine_identity_at_size():
    """Size-independent property at 110,592 quadrature points and 5,000
    targets: with the fine level equal to the coarse one (each point its own
    child, no truncation), psi_f = psi_c - lambda log(1 + g_c / nu_c), g from
    the FP64 evaluation at the same potential with epsilon = lambda."""
    say("refine identity at 24^3 x 5000")
    src, tgt = structured_case(3, 24, 2, 5000, 41, gmm_source=True, gmm_targets=True,
                               dirichlet=5.0)
    lam = 0.01
    s = solver_for(src)
    s.setup_target(tgt.y, tgt.nu)
    s.configure(sd.SolverParameters(epsilon=lam, tau=0.0, use_log_sum_exp_trick=True))
    psi = 0.02 * np.random.default_rng(41).standard_normal(tgt.n)
    _, g, _ = s.evaluate_functional_R(DMAX, psi)
    pg = refine_gpu(s, tgt.y, tgt, psi, [[i] for i in range(tgt.n)], lam, DMAX, True)
    pred = psi - lam * np.log1p(g / tgt.nu)
    # g_j = nu_j m_j - nu_j carries an absolute error of about one rounding of
    # nu_j, so log1p(g_j / nu_j) = log(m_j) is known to ~1e-16 / m_j where the
    # target's mass ratio m_j = 1 + g_j / nu_j is small
    m = 1.0 + g / tgt.nu
    tol = lam * (1e-11 + 4e-16 / m) + 1e-13 * np.abs(psi)
    assert np.all(np.abs(pg - pred) <= tol), float(np.max(np.abs(pg - pred) / tol))


def test_refine_argument_errors():
    src, _ = structured_case(2, 4, 2, 4, 1)
    s = solver_for(src)
    r = sd.SoftmaxRefinement(s, 0.5, True)
    yc = np.random.default_rng(0).random((3, 2))
    with pytest.raises(sd.SdotError):
        r.compute_refinement(yc, None, yc, np.ones(3) / 3, np.zeros(3), 0.0, 0, [[[0], [1], [2]]])
    # an invalid level: every fine potential 0 (the reference reports and returns)
    out = r.compute_refinement(yc, None, yc, np.ones(3) / 3, np.zeros(3), 0.01, 2, [[[0], [1], [2]]])
    assert np.all(out == 0.0)


def test_target_multilevel_driver(tmp_path):
    """run_target_multilevel over a 3-level hierarchy (FP64, no truncation):
    every level converges, and the finest potential solves the finest problem
    (componentwise marginal error within the tolerance at an independent
    evaluation)."""
    say("target multilevel driver")
    src_o, _ = structured_case(2, 24, 2, 4, 51, gmm_source=True)
    rng = np.random.default_rng(51)
    y = gmm_sample(rng, 600, 2, 4)
    nu = rng.dirichlet(np.full(600, 4.0))
    hdir = str(tmp_path / "hier")
    sd.generate_target_hierarchy(y, nu, hdir, sizes=[600, 120, 30], seed=2)
    s = solver_for(src_o)
    prm = sd.SolverParameters(epsilon=0.01, tau=0.0, tolerance=1e-7,
                              solver_control_type="componentwise", max_iterations=2000,
                              use_log_sum_exp_trick=True, precision="fp64",
                              line_search="noise_aware")
    ml = sd.MultilevelParameters(target_hierarchy_dir=hdir)
    psi, recs = sd.run_target_multilevel(s, prm, ml, output_dir=str(tmp_path / "out"))
    assert [r.level for r in recs] == [2, 1, 0]
    assert all(r.status == "converged" for r in recs), [(r.level, r.status) for r in recs]
    assert os.path.exists(tmp_path / "out" / "potentials.txt")
    yl = sd.formats.read_vector(os.path.join(hdir, "level_0_points"), dim=2)
    nl = sd.formats.read_vector(os.path.join(hdir, "level_0_density"))
    s.setup_target(yl.ravel(), nl)
    s.configure(prm)
    _, g, _ = s.evaluate_functional_R(DMAX, psi)
    assert np.max(np.abs(g) / nl) <= 1e-7 * 1.001


def test_combined_multilevel_driver(tmp_path):
    """run_combined_multilevel over two source levels (8^3, 16^3 hexes) and
    three target levels, epsilon 0.09 -> 0.03 -> 0.01 spread over the
    combined levels, FP32_LSE solves: the source stays at its coarsest level
    until the target hierarchy has as many levels left, every level reaches
    the tolerance, and the final potential solves the finest problem in FP64
    within 1e-6 relative marginal error."""
    say("combined multilevel driver")
    rng = np.random.default_rng(61)
    srcs = []
    for n in (8, 16):
        src_o, _ = structured_case(3, n, 2, 4, 61, gmm_source=True)
        srcs.append(sd.SourceArrays(3, src_o.nq, src_o.x, src_o.w, src_o.center))
    y = gmm_sample(rng, 1500, 3, 5)
    nu = rng.dirichlet(np.full(1500, 5.0))
    hdir = str(tmp_path / "hier")
    sd.generate_target_hierarchy(y, nu, hdir, sizes=[1500, 300, 60], seed=2)
    s = sd.SotSolver(3)
    prm = sd.SolverParameters(epsilon=0.01, tau=0.0, tolerance=1e-6,
                              solver_control_type="componentwise", max_iterations=3000,
                              use_log_sum_exp_trick=True, precision="fp32_lse",
                              line_search="noise_aware", use_epsilon_scaling=True,
                              epsilon_scaling_factor=3.0, epsilon_scaling_steps=3)
    ml = sd.MultilevelParameters(target_hierarchy_dir=hdir)
    psi, recs = sd.run_combined_multilevel(s, srcs, prm, ml)
    assert [(r.level, r.n_cells) for r in recs] == [(2, 512), (1, 512), (0, 4096)]
    assert [r.epsilons for r in recs] == [[0.09], [0.03], [0.01]]
    assert all(r.status == "converged" for r in recs), [(r.level, r.status) for r in recs]
    s.setup_target(y.ravel(), nu)
    s.configure(sd.SolverParameters(epsilon=0.01, tau=0.0, use_log_sum_exp_trick=True))
    _, g, _ = s.evaluate_functional_R(DMAX, psi)
    assert np.max(np.abs(g) / nu) <= 1e-6 * 1.001
