"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle and
the 50-digit golden vectors.

Tolerances (FP64 mode): |F - F_ref| <= 1e-10 |F_ref| and
max_j |g_j - g_ref_j| <= 1e-10 * max_j (g_ref_j + nu_j), i.e. relative to the
assembled marginal; candidate sets per cell identical (count and hash).
FP32_LSE mode: relative marginal error max_j |g_j - g_ref_j| / nu_j <= 1e-6
and |F - F_ref| <= 1e-8 |F_ref|."""
import json
import os
import sys

import numpy as np
import pytest

import semidiscreteot_amd as sd
from oracle import oracle as O
from tests.cases import structured_case, tet_case

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "tiny_cases.json")))
DMAX = np.finfo(float).max


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def make(src, tgt, **kw):
    s = sd.SotSolver(src.dim)
    s.setup_source(src.x, src.w, src.center, src.nq)
    s.setup_target(tgt.y, tgt.nu)
    s.configure(sd.SolverParameters(**kw))
    return s


def assert_fp64(F, g, Fo, go, nu, tol=1e-10):
    assert abs(F - Fo) <= tol * abs(Fo), (F, Fo)
    scale = np.max(np.abs(go + nu))
    err = np.max(np.abs(g - go))
    assert err <= tol * scale, (err, scale)


@pytest.mark.parametrize("case", GOLD, ids=[c["name"] for c in GOLD])
def test_golden_fp64(case):
    say(f"golden {case['name']}")
    src = O.Source(case["dim"], case["nq"], case["x"], case["rho"], case["jxw"], case["center"])
    tgt = O.Target(case["y"], case["nu"])
    s = make(src, tgt, epsilon=case["eps"], use_log_sum_exp_trick=case["use_lse"])
    F, g, C = s.evaluate_functional_R(case["R"], case["psi"])
    gg = np.array([float(v) for v in case["g"]])
    assert_fp64(F, g, float(case["F"]), gg, np.array(case["nu"]), tol=1e-12)
    assert abs(C - float(case["C"])) <= 1e-12 * float(case["C"])
    cnt, h, _ = s.candidate_stats(case["R"])
    assert cnt.tolist() == case["count"]
    assert [str(v) for v in h.tolist()] == case["hash"]


@pytest.mark.parametrize("lse", [True, False])
def test_untruncated_2d(lse):
    say(f"untruncated 2d lse={lse}")
    src, tgt = structured_case(2, 32, 3, 400, seed=0)
    psi = 0.01 * np.random.default_rng(0).standard_normal(tgt.n)
    s = make(src, tgt, epsilon=1e-2, use_log_sum_exp_trick=lse)
    F, g, C = s.evaluate_functional_R(DMAX, psi)
    Fo, go, Co = O.evaluate_R(src, tgt, O.params(epsilon=1e-2, use_lse=lse), DMAX, psi)
    assert_fp64(F, g, Fo, go, tgt.nu)
    assert abs(C - Co) <= 1e-10 * Co


def test_plain_mode_underflow_skips_q():
    say("plain underflow")
    src, tgt = structured_case(2, 12, 2, 5, seed=9)
    y = np.full((5, 2), 0.03) + 0.01 * np.arange(5)[:, None]
    tgt = O.Target(y.ravel(), np.full(5, 0.2))
    s = make(src, tgt, epsilon=2e-4, use_log_sum_exp_trick=False)
    psi = np.zeros(5)
    F, g, C = s.evaluate_functional_R(DMAX, psi)
    Fo, go, Co = O.evaluate_R(src, tgt, O.params(epsilon=2e-4, use_lse=0), DMAX, psi)
    assert Co > 0
    # q-points with a subnormal sum give w/S = inf and an infinite gradient entry in
    # the reference arithmetic; the same entries must be infinite here
    inf = ~np.isfinite(go)
    assert np.array_equal(inf, ~np.isfinite(g)) and np.array_equal(g[inf], go[inf])
    fin = ~inf
    if np.isfinite(Fo):
        assert abs(F - Fo) <= 1e-10 * abs(Fo)
    if fin.any():
        assert np.max(np.abs(g[fin] - go[fin])) <= 1e-10 * np.max(np.abs(go[fin] + tgt.nu[fin]))


@pytest.mark.parametrize("precision", ["fp64", "fp32_lse"])
def test_truncated_3d_sets_and_values(precision):
    say(f"truncated 3d {precision}")
    src, tgt = structured_case(3, 10, 2, 2000, seed=1, gmm_source=True, gmm_targets=True,
                               dirichlet=5)
    rng = np.random.default_rng(2)
    psi = 0.01 * rng.standard_normal(tgt.n)
    eps, tau = 1e-2, 1e-10
    st = O.new_state()
    R = O.distance_threshold(src, tgt, O.params(epsilon=eps, tau=tau), st, psi)
    assert R < 1.0  # truncated, not dense
    s = make(src, tgt, epsilon=eps, tau=tau, use_log_sum_exp_trick=True, precision=precision)
    assert s.distance_threshold(psi) == R
    cnt, h, pairs = s.candidate_stats(R)
    cnt_o, h_o = O.candidate_stats(src, tgt, R)
    assert np.array_equal(cnt, cnt_o) and np.array_equal(h, h_o)
    assert 0 < pairs < src.n_cells * src.nq * tgt.n
    F, g, C = s.evaluate_functional_R(R, psi)
    Fo, go, Co = O.evaluate_R(src, tgt, O.params(epsilon=eps, use_lse=1), R, psi)
    if precision == "fp64":
        assert_fp64(F, g, Fo, go, tgt.nu)
    else:
        assert abs(F - Fo) <= 1e-8 * abs(Fo)
        assert np.max(np.abs(g - go) / tgt.nu) <= 1e-6
        # C_global relative to each q-point's largest term, as the reference sums it
        assert abs(C - Co) <= 1e-6 * Co


def test_fp32_small_eps_error():
    say("fp32 eps=1e-3")
    src, tgt = structured_case(3, 12, 2, 3000, seed=5, gmm_source=True, gmm_targets=True,
                               dirichlet=5)
    psi = 0.02 * np.random.default_rng(3).standard_normal(tgt.n)
    s = make(src, tgt, epsilon=1e-3, use_log_sum_exp_trick=True, precision="fp32_lse")
    F, g, C = s.evaluate_functional_R(DMAX, psi)
    Fo, go, Co = O.evaluate_R(src, tgt, O.params(epsilon=1e-3, use_lse=1), DMAX, psi)
    rel = np.max(np.abs(g - go) / tgt.nu)
    say(f"fp32 rel marginal err {rel:.3e}  F rel {abs(F - Fo) / abs(Fo):.3e}")
    assert rel <= 1e-6
    assert abs(F - Fo) <= 1e-8 * abs(Fo)
    # the next evaluations start their online sweep from the previous maxima
    # (history seed): a small psi change is seeded, a large one is not
    # (psi is far from the solution here: a target's mass m_j can be orders of
    # magnitude above nu_j, and the FP32 error is relative to the mass)
    for step in (2e-4, 3e-2):
        psi2 = psi + step * np.random.default_rng(4).standard_normal(tgt.n)
        F, g, C = s.evaluate_functional_R(DMAX, psi2)
        Fo, go, Co = O.evaluate_R(src, tgt, O.params(epsilon=1e-3, use_lse=1), DMAX, psi2)
        rel = np.max(np.abs(g - go) / tgt.nu)
        relm = np.max(np.abs(g - go) / np.maximum(tgt.nu, go + tgt.nu))
        say(f"fp32 next eval (dpsi {step:g}): error / nu {rel:.3e}, error / max(nu, m) {relm:.3e}, "
            f"max m/nu {np.max((go + tgt.nu) / tgt.nu):.3g}")
        assert relm <= 1e-6
        assert abs(F - Fo) <= 1e-8 * abs(Fo)
        psi = psi2


@pytest.mark.parametrize("prune", [0.0, 50.0])
def test_fp32_seed_across_radius_changes(prune):
    """FP32 sweeps are seeded with the previous evaluation's maxima when the
    radius changed: a larger radius adds candidates the seed does not bound
    (their items are swept again, among them points that had no candidate at
    all), a smaller one removes some (value-bounded sweeps: not seeded, since
    their bound rests on the previous candidates).  Every evaluation of the
    sequence matches the oracle at the FP32 bar."""
    say(f"fp32 seed across radii, prune_bits {prune}")
    src, tgt = structured_case(3, 12, 2, 3000, seed=5, gmm_source=True, gmm_targets=True,
                               dirichlet=5)
    eps = 1e-3
    rng = np.random.default_rng(21)
    psi = 0.02 * rng.standard_normal(tgt.n)
    s = make(src, tgt, epsilon=eps, use_log_sum_exp_trick=True, precision="fp32_lse",
             prune_bits=prune)
    for R, step in ((0.08, 0.0), (0.3, 2e-4), (0.2, 2e-4), (0.35, 3e-2), (0.1, 2e-4), (0.3, 0.0),
                    (0.32, 2e-4), (0.34, 2e-4)):
        psi = psi + step * rng.standard_normal(tgt.n)
        F, g, C = s.evaluate_functional_R(R, psi)
        Fo, go, Co = O.evaluate_R(src, tgt, O.params(epsilon=eps, use_lse=1), R, psi)
        relm = np.max(np.abs(g - go) / np.maximum(tgt.nu, go + tgt.nu))
        say(f"R {R}: error / max(nu, m) {relm:.3e}, F rel {abs(F - Fo) / abs(Fo):.3e}")
        assert relm <= 1e-6
        assert abs(F - Fo) <= 1e-8 * abs(Fo)


@pytest.mark.parametrize("precision", ["fp64", "fp32_lse"])
def test_stateful_threshold_sequence(precision):
    """evaluate_functional reads C_global / the previous functional for the
    integral and geometric bounds (SotSolver.cc:392-404).  FP32: the state it
    carries (C_global relative to each q-point's largest term, F) is within
    1e-6 / 1e-8 of the oracle's, so the thresholds agree to 1e-9."""
    say(f"stateful thresholds {precision}")
    src, tgt = structured_case(2, 16, 2, 300, seed=7, dirichlet=5)
    rng = np.random.default_rng(7)
    for ttype in ("pointwise", "integral", "geometric"):
        s = make(src, tgt, epsilon=1e-2, tau=1e-9, use_log_sum_exp_trick=True,
                 distance_threshold_type=ttype, precision=precision)
        prm = O.params(epsilon=1e-2, tau=1e-9, use_lse=1, threshold=ttype)
        st = O.new_state()
        for k in range(3):
            psi = 0.02 * rng.standard_normal(tgt.n)
            F, g = s.evaluate_functional(psi)
            Fo, go = O.evaluate(src, tgt, prm, st, psi)
            if precision == "fp64":
                assert s.get_last_distance_threshold() == st.last_threshold, ttype
                assert_fp64(F, g, Fo, go, tgt.nu)
                assert abs(s.get_C_global() - st.C_global) <= 1e-10 * st.C_global
            else:
                R, Ro = s.get_last_distance_threshold(), st.last_threshold
                assert abs(R - Ro) <= 1e-9 * Ro, (ttype, R, Ro)
                assert abs(F - Fo) <= 1e-8 * abs(Fo)
                assert np.max(np.abs(g - go) / np.maximum(tgt.nu, go + tgt.nu)) <= 1e-6
                assert abs(s.get_C_global() - st.C_global) <= 1e-6 * st.C_global


def test_covering_radius_exact():
    say("covering radius")
    src, tgt = structured_case(3, 8, 2, 500, seed=4, gmm_targets=True)
    s = make(src, tgt, epsilon=1e-2)
    assert s.compute_covering_radius() == O.covering_radius(src, tgt)
    # kept per source and target set: new targets, then a new source
    src2, tgt2 = structured_case(3, 6, 3, 300, seed=5, gmm_targets=True)
    s.setup_target(tgt2.y, tgt2.nu)
    assert s.compute_covering_radius() == O.covering_radius(src, tgt2)
    s.setup_source(src2.x, src2.w, src2.center, src2.nq)
    assert s.compute_covering_radius() == O.covering_radius(src2, tgt2)
    assert s.compute_covering_radius() == O.covering_radius(src2, tgt2)
    s.close()


def test_determinism_bitwise():
    say("determinism")
    src, tgt = structured_case(3, 10, 2, 1500, seed=6, gmm_targets=True)
    s = make(src, tgt, epsilon=5e-3, tau=1e-10, use_log_sum_exp_trick=True, precision="fp32_lse")
    psi = 0.01 * np.random.default_rng(6).standard_normal(tgt.n)
    R = s.distance_threshold(psi)
    # FP32 mode seeds each sweep with the previous evaluation's reference values, so
    # the result is a deterministic function of the call sequence: the first call
    # (unseeded) may differ in the last bits from the later (seeded) ones, which are
    # identical to each other
    a = s.evaluate_functional_R(R, psi)
    b = s.evaluate_functional_R(R, psi)
    c = s.evaluate_functional_R(R, psi)
    assert b[0] == c[0] and np.array_equal(b[1], c[1]) and b[2] == c[2]
    assert abs(a[0] - b[0]) <= 1e-9 * abs(b[0])
    s.configure(sd.SolverParameters(epsilon=5e-3, tau=1e-10, use_log_sum_exp_trick=True))
    a = s.evaluate_functional_R(R, psi)
    b = s.evaluate_functional_R(R, psi)
    assert a[0] == b[0] and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_tet_source_parity():
    say("tets")
    src, tgt, _ = tet_case(2000, 400, seed=3)
    psi = 0.01 * np.random.default_rng(8).standard_normal(tgt.n)
    s = make(src, tgt, epsilon=5e-3, tau=1e-10, use_log_sum_exp_trick=True)
    R = s.distance_threshold(psi)
    F, g, C = s.evaluate_functional_R(R, psi)
    Fo, go, Co = O.evaluate_R(src, tgt, O.params(epsilon=5e-3, use_lse=1), R, psi)
    assert_fp64(F, g, Fo, go, tgt.nu)
    cnt, h, _ = s.candidate_stats(R)
    cnt_o, h_o = O.candidate_stats(src, tgt, R)
    assert np.array_equal(cnt, cnt_o) and np.array_equal(h, h_o)


def test_edge_cases():
    say("edge cases")
    src, _ = structured_case(2, 8, 3, 1, seed=0)
    # single target: closed form
    tgt = O.Target(np.array([0.3, 0.6]), np.array([1.0]))
    s = make(src, tgt, epsilon=1e-2, use_log_sum_exp_trick=True)
    F, g, C = s.evaluate_functional_R(DMAX, np.zeros(1))
    assert abs(F + 0.5 * (0.2 ** 2 + 0.1 ** 2 + 1 / 6)) < 1e-13 and abs(g[0]) < 1e-13
    # empty candidate sets everywhere: F = -psi.nu, g = -nu
    rng = np.random.default_rng(1)
    tgt = O.Target(rng.random(40) + 5.0, np.full(20, 0.05))
    s = make(src, tgt, epsilon=1e-2, use_log_sum_exp_trick=True)
    psi = rng.standard_normal(20)
    F, g, C = s.evaluate_functional_R(0.5, psi)
    assert abs(F + psi @ tgt.nu) < 1e-15 and np.allclose(g, -tgt.nu) and C == 0.0
    Fo, go, _ = O.evaluate_R(src, tgt, O.params(epsilon=1e-2, use_lse=1), 0.5, psi)
    assert abs(F - Fo) < 1e-15
    # NaN threshold: no candidates (sqrt(s) <= NaN is false)
    F, g, C = s.evaluate_functional_R(float("nan"), psi)
    assert np.allclose(g, -tgt.nu)


def test_errors_raise():
    say("errors")
    s = sd.SotSolver(2)
    with pytest.raises(sd.SdotError):
        s.evaluate_functional(np.zeros(3))
    src, tgt = structured_case(2, 4, 2, 5, seed=0)
    s.setup_source(src.x, src.w, src.center, src.nq)
    with pytest.raises(ValueError):
        s.setup_target(np.zeros(7), np.ones(5))
    with pytest.raises(sd.SdotError):
        s.configure(sd.SolverParameters(epsilon=0.0))


def test_solve_matches_oracle():
    say("solve")
    src, tgt = structured_case(2, 16, 3, 60, seed=12, gmm_source=True, dirichlet=5)
    kw = dict(epsilon=1e-2, tau=1e-12, tolerance=1e-6, solver_control_type="componentwise",
              max_iterations=300, use_log_sum_exp_trick=True)
    s = make(src, tgt, **kw)
    psi = s.solve()
    info = s.get_last_solve_info()
    prm = O.params(epsilon=1e-2, tau=1e-12, tolerance=1e-6, control="componentwise",
                   max_iterations=300, use_lse=1)
    psi_o, info_o = O.solve(src, tgt, prm)
    say(f"gpu {info}  oracle {info_o['status']} it={info_o['iterations']}")
    assert info["status"] == info_o["status"] == "converged"
    assert s.get_convergence_status()
    assert abs(info["iterations"] - info_o["iterations"]) <= 2
    assert np.max(np.abs((psi - psi.mean()) - (psi_o - psi_o.mean()))) < 1e-6
    Fo, go, _ = O.evaluate_R(src, tgt, prm, DMAX, psi)
    assert np.max(np.abs(go) / tgt.nu) <= 1e-6 * 1.001


@pytest.mark.parametrize("dim,tau", [(2, 0.0), (3, 1e-6)])
def test_hessian_matches_oracle(dim, tau):
    """compute_hessian (declared at SotSolver.h:338-344): GPU MFMA FP64 SYRK
    against the oracle's dense assembly; H*1 = 0 and symmetry."""
    say(f"hessian dim={dim} tau={tau}")
    if dim == 2:
        src, tgt = structured_case(2, 16, 3, 150, seed=21)
    else:
        src, tgt = structured_case(3, 6, 2, 200, seed=22, gmm_targets=True, dirichlet=5)
    psi = 0.01 * np.random.default_rng(4).standard_normal(tgt.n)
    s = make(src, tgt, epsilon=2e-2, tau=tau, use_log_sum_exp_trick=True)
    R = s.distance_threshold(psi)
    if dim == 3:
        assert R < 1.0
    H = s.compute_hessian(psi)
    Ho = O.hessian(src, tgt, O.params(epsilon=2e-2, use_lse=1), R, psi)
    scale = np.max(np.abs(Ho))
    assert np.max(np.abs(H - Ho)) <= 1e-10 * scale
    assert np.max(np.abs(H - H.T)) <= 1e-12 * scale
    assert np.max(np.abs(H.sum(1))) <= 1e-10 * scale


@pytest.mark.parametrize("tau", [0.0, 1e-6])
def test_hessian_large_n_matches_oracle(tau):
    """compute_hessian at N = 2,050 (17 SYRK column blocks of 128, the last
    one padded; 4,096 q-points in one chunk of 16-row stages) against the
    oracle's assembly, untruncated and truncated."""
    say(f"hessian N=2050 tau={tau}")
    src, tgt = structured_case(3, 8, 2, 2050, seed=23, gmm_source=True, gmm_targets=True,
                               dirichlet=5)
    psi = 0.005 * np.random.default_rng(5).standard_normal(tgt.n)
    s = make(src, tgt, epsilon=1e-2, tau=tau, use_log_sum_exp_trick=True)
    R = s.distance_threshold(psi)
    H = s.compute_hessian(psi)
    Ho = O.hessian(src, tgt, O.params(epsilon=1e-2, use_lse=1), R, psi)
    scale = np.max(np.abs(Ho))
    err = np.max(np.abs(H - Ho)) / scale
    say(f"R {R:.4g}: max |H - H_ref| / max |H_ref| = {err:.3e}")
    assert err <= 1e-10
    assert np.max(np.abs(H - H.T)) <= 1e-12 * scale
    assert np.max(np.abs(H.sum(1))) <= 1e-10 * scale


def test_discarded_mass_audit():
    """sdot_q_log_partition: the discarded mass of a radius matches the oracle's
    direct count on sampled quadrature points."""
    say("discarded mass")
    src, tgt = structured_case(3, 8, 2, 800, seed=31, gmm_targets=True, dirichlet=5)
    psi = 0.02 * np.random.default_rng(2).standard_normal(tgt.n)
    s = make(src, tgt, epsilon=5e-3, use_log_sum_exp_trick=True)
    R = 0.25
    lz_full = s.q_log_partition(DMAX, psi)
    lz = s.q_log_partition(R, psi)
    frac = -np.expm1(np.minimum(lz - lz_full, 0.0))
    q = np.arange(0, src.n_cells * src.nq, 37)
    fo = O.discarded_fraction(src, tgt, O.params(epsilon=5e-3, use_lse=1), R, psi, q)
    assert np.max(np.abs(frac[q] - fo)) < 1e-12
    assert abs(s.discarded_mass(R, psi, src.w) - float(np.sum(src.w * frac))) < 1e-15


def test_plain_2d_config_solve():
    """The reference's CPU-runnable configuration end to end: 64x64 Q1 / QGauss(3)
    on the unit square (36,864 q), 1,000 uniform targets (seed 0), eps = 1e-2,
    no truncation, FP64, l1-norm tolerance 1e-8 from psi = 0.  With the
    reference's line search the GPU solve takes the oracle's decisions for the
    first 60 iterations (same evaluation count, potentials and functional to
    rounding).  Near ||g||_1 ~ 5e-8 the functional's rounding noise decides
    the reference search on both sides (the oracle stops there with a failed
    line search); the noise-aware search reaches the tolerance, confirmed on
    the oracle's gradient."""
    say("plain 2D configuration")
    src = O.structured_source(2, [64, 64], [0.0, 0.0], [1.0, 1.0], 3)
    y = np.random.default_rng(0).random((1000, 2))
    tgt = O.Target(y.ravel(), np.full(1000, 1e-3))
    kw = dict(epsilon=1e-2, tau=0.0, tolerance=1e-8, solver_control_type="l1norm",
              use_log_sum_exp_trick=True, verbose_output=False)
    s = make(src, tgt, max_iterations=60, **kw)
    psi = s.solve()
    info = s.get_last_solve_info()
    prm = O.params(epsilon=1e-2, tau=0.0, tolerance=1e-8, control="l1norm", max_iterations=60,
                   use_lse=1)
    psi_o, info_o = O.solve(src, tgt, prm)
    say(f"60 its: gpu {info['status']} ev={info['evaluations']} cv={info['check_value']:.3e}; "
        f"oracle {info_o['status']} ev={info_o['evaluations']} cv={info_o['check_value']:.3e}")
    assert info["status"] == info_o["status"] == "max_iterations"
    assert info["evaluations"] == info_o["evaluations"]
    assert np.max(np.abs(psi - psi_o)) < 1e-9
    assert abs(info["functional"] - info_o["functional"]) <= 1e-13 * abs(info_o["functional"])
    s.configure(sd.SolverParameters(line_search="noise_aware", max_iterations=1000, **kw))
    psi2 = s.solve()
    info2 = s.get_last_solve_info()
    _, g, _ = O.evaluate_R(src, tgt, prm, DMAX, psi2)
    say(f"noise-aware: {info2['status']} it={info2['iterations']} oracle ||g||_1 {np.abs(g).sum():.3e}")
    assert info2["status"] == "converged" and np.abs(g).sum() <= 1e-8 * 1.01
    s.close()


@pytest.mark.parametrize("precond,precision,control", [
    ("none", "fp64", "l1norm"), ("diagonal", "fp64", "l1norm"),
    ("diagonal", "fp64", "componentwise"), ("diagonal", "fp32_lse", "l1norm")])
def test_device_lbfgs(precond, precision, control, monkeypatch):
    """Solves of 32,768 targets and more keep every L-BFGS vector in HBM (iterate,
    gradient, direction, pairs) and run the vector arithmetic there (dot
    products and stopping statistics over a fixed grid of partial sums).
    Forced on (SDOT_LBFGS_DEVICE_MIN=0) for the plain 2D configuration
    (1,000 targets, 60 iterations from psi = 0), the solve
    takes the decisions of the host driver: same status and evaluation count,
    potentials and functional to rounding (FP64); with FP32 evaluations and
    the FP64 finish, the same status and potentials within 1e-9 of their
    range."""
    say(f"device L-BFGS, {precond} {precision} {control}")
    src = O.structured_source(2, [64, 64], [0.0, 0.0], [1.0, 1.0], 3)
    y = np.random.default_rng(0).random((1000, 2))
    tgt = O.Target(y.ravel(), np.full(1000, 1e-3))
    # componentwise: max_j |g_j| - 1e-5 nu_j, reached before the functional's
    # rounding decides the reference line search
    tol = 1e-8 if control == "l1norm" else 1e-5
    s = make(src, tgt, epsilon=1e-2, tau=0.0, tolerance=tol, solver_control_type=control,
             use_log_sum_exp_trick=True, verbose_output=False, max_iterations=60,
             preconditioner=precond, precision=precision)
    monkeypatch.delenv("SDOT_LBFGS_DEVICE_MIN", raising=False)
    psi_h = s.solve()
    info_h = s.get_last_solve_info()
    monkeypatch.setenv("SDOT_LBFGS_DEVICE_MIN", "0")
    psi_d = s.solve()
    info_d = s.get_last_solve_info()
    monkeypatch.delenv("SDOT_LBFGS_DEVICE_MIN")
    d = float(np.max(np.abs(psi_d - psi_h)))
    say(f"host {info_h['status']} ev={info_h['evaluations']} F={info_h['functional']:.15e}; "
        f"device {info_d['status']} ev={info_d['evaluations']} F={info_d['functional']:.15e}; "
        f"max |dpsi| {d:.2e}")
    assert info_d["status"] == info_h["status"]
    if precision == "fp64":
        assert info_d["evaluations"] == info_h["evaluations"]
        assert d < 1e-9
        assert abs(info_d["functional"] - info_h["functional"]) <= 1e-13 * abs(info_h["functional"])
    else:
        assert d <= 1e-9 * float(np.ptp(psi_h))
    s.close()


@pytest.mark.parametrize("control", ["l1norm", "componentwise"])
def test_device_lbfgs_iteration_lines(control, capfd, monkeypatch):
    """The iteration lines of VerboseSolverControl::check (SotSolver.h:408-440)
    from the statistics kernel of the device driver (L-2 norm, L-1 norm or the
    componentwise residual, printed to 6 digits) match the host driver's line
    for line on 12 FP64 iterations of the plain 2D configuration."""
    say(f"device L-BFGS iteration lines, {control}")
    src = O.structured_source(2, [64, 64], [0.0, 0.0], [1.0, 1.0], 3)
    y = np.random.default_rng(0).random((1000, 2))
    tgt = O.Target(y.ravel(), np.full(1000, 1e-3))
    s = make(src, tgt, epsilon=1e-2, tau=0.0, tolerance=1e-8, solver_control_type=control,
             use_log_sum_exp_trick=True, verbose_output=True, max_iterations=12,
             preconditioner="diagonal")
    lines = {}
    for mode in ("host", "device"):
        if mode == "device":
            monkeypatch.setenv("SDOT_LBFGS_DEVICE_MIN", "0")
        else:
            monkeypatch.delenv("SDOT_LBFGS_DEVICE_MIN", raising=False)
        capfd.readouterr()
        s.solve()
        out = capfd.readouterr().out
        lines[mode] = [ln for ln in out.splitlines() if ln.startswith("Iteration ")]
    monkeypatch.delenv("SDOT_LBFGS_DEVICE_MIN")
    s.close()
    say(f"{len(lines['host'])} lines; last: {lines['device'][-1]}")
    assert len(lines["host"]) == 13
    assert lines["device"] == lines["host"]


@pytest.mark.parametrize("which", ["targets", "large_source"])
def test_large_record_sets_match_oracle(which):
    """Record sets larger than an XCD's L2: 70,000 targets (4.5 MB of target
    records for the q-major sweeps), truncated; and a large source (551,368
    quadrature points, whose q records outgrow the L2 of the target-major
    sweep).  FP64 to 1e-10 (bitwise repeatable) and FP32 to 1e-6 of each
    target's mass against the oracle, barycenters in FP64 to 1e-10."""
    say(f"large record sets: {which}")
    if which == "targets":
        src, tgt = structured_case(3, 8, 2, 70000, seed=51, gmm_source=True, gmm_targets=True,
                                   dirichlet=5)
        eps, R = 1e-2, 0.15
    else:
        src, tgt = structured_case(3, 41, 2, 3000, seed=52, gmm_source=True, gmm_targets=True,
                                   dirichlet=5)
        eps, R = 5e-3, 0.1
    psi = 0.002 * np.random.default_rng(5).standard_normal(tgt.n)
    prm = O.params(epsilon=eps, use_lse=1)
    Fo, go, Co = O.evaluate_R(src, tgt, prm, R, psi)
    for precision in ("fp64", "fp32_lse"):
        s = make(src, tgt, epsilon=eps, use_log_sum_exp_trick=True, precision=precision)
        F, g, C = s.evaluate_functional_R(R, psi)
        F2, g2, _ = s.evaluate_functional_R(R, psi)   # seeded FP32 evaluation
        if precision == "fp64":
            assert_fp64(F, g, Fo, go, tgt.nu)
            assert F2 == F and np.array_equal(g2, g)
            b = s.compute_weighted_barycenters_R(R, psi)
            bo = O.barycenters_R(src, tgt, prm, R, psi)
            assert np.max(np.abs(b - bo)) <= 1e-10
        else:
            # away from the solution a target's mass m_j can be many times nu_j;
            # the FP32 error is relative to the mass (m = nu at the solution)
            scale = np.maximum(tgt.nu, go + tgt.nu)
            for FF, gg in ((F, g), (F2, g2)):
                assert abs(FF - Fo) <= 1e-8 * abs(Fo)
                assert np.max(np.abs(gg - go) / scale) <= 1e-6
        s.close()
