"""Segment counts and work-item sizes the sizing rules pick only at sizes no
other test reaches, checked against the FP64 oracle.

  * single-solve evaluations with the q-major and target-major segment
    counts fixed through the SDOT_TEST_SEGMENTS hook (K1 in {1, 3, 7, 32},
    K2 in {1, 5, 8, 128}; counts that are not multiples of 8 leave the
    per-XCD row split), FP64 and FP32_LSE, dense and truncated;
  * served solves with small work items (8 targets per q-major item and 32
    q-points per target-major item: about 100 segments per pass) against the
    FP64 oracle and against the default item sizes.

Tolerances as in test_gpu_parity.py: FP64 |F - F_ref| <= 1e-10 |F_ref|,
max_j |g_j - g_ref_j| <= 1e-10 max_j (g_ref_j + nu_j); FP32_LSE F within 1e-8
and the gradient within 1e-6 of max(nu_j, m_j)."""
import os
import sys

import numpy as np
import pytest

import semidiscreteot_amd as sd
from oracle import oracle as O
from tests.cases import gmm_sample, structured_case

pytestmark = pytest.mark.gpu
DMAX = np.finfo(float).max


def say(msg):
    print(f"[gpu-seg] {msg}", flush=True)
    sys.stdout.flush()


def solver_with_segments(k1, k2, dim):
    os.environ["SDOT_TEST_SEGMENTS"] = f"{k1},{k2}"
    try:
        return sd.SotSolver(dim)
    finally:
        del os.environ["SDOT_TEST_SEGMENTS"]


@pytest.mark.parametrize("k1,k2", [(1, 1), (3, 5), (32, 8), (7, 128), (1, 128)])
@pytest.mark.parametrize("precision", ["fp64", "fp32_lse"])
def test_fixed_segment_counts_vs_oracle(k1, k2, precision):
    cases = [("3d truncated", structured_case(3, 6, 2, 700, seed=3, gmm_source=True,
                                              gmm_targets=True, dirichlet=5), 5e-3, 0.35),
             ("2d dense", structured_case(2, 10, 3, 300, seed=4, gmm_source=True), 1e-2, DMAX)]
    for name, (src, tgt), eps, R in cases:
        say(f"K1={k1} K2={k2} {precision} {name}")
        s = solver_with_segments(k1, k2, src.dim)
        s.setup_source(src.x, src.w, src.center, src.nq)
        s.setup_target(tgt.y, tgt.nu)
        s.configure(sd.SolverParameters(epsilon=eps, use_log_sum_exp_trick=True,
                                        precision=precision))
        psi = 0.01 * np.random.default_rng(k1 * 100 + k2).standard_normal(tgt.n)
        Fo, go, _ = O.evaluate_R(src, tgt, O.params(epsilon=eps, use_lse=1), R, psi)
        # two evaluations: the second one is seeded in FP32_LSE
        for psi_e in (psi, psi + 1e-4 * np.random.default_rng(1).standard_normal(tgt.n)):
            F, g, _ = s.evaluate_functional_R(R, psi_e)
            Fo, go, _ = O.evaluate_R(src, tgt, O.params(epsilon=eps, use_lse=1), R, psi_e)
            if precision == "fp64":
                assert abs(F - Fo) <= 1e-10 * abs(Fo)
                assert np.max(np.abs(g - go)) <= 1e-10 * np.max(np.abs(go + tgt.nu))
            else:
                assert abs(F - Fo) <= 1e-8 * abs(Fo)
                rel = np.max(np.abs(g - go) / np.maximum(tgt.nu, go + tgt.nu))
                assert rel <= 1e-6, rel
        s.close()


def test_served_small_work_items_vs_oracle_and_default_sizes():
    """About 100-128 segments per pass (K1 = N / 8 capped at 128, K2 =
    nQ / 32 = 128 on the 8^3 order-2 mesh): the served solves meet the
    tolerance under the FP64 oracle and agree with the default-size solves."""
    src, _ = structured_case(3, 8, 2, 16, seed=1, gmm_source=True)
    s = sd.SotSolver(3)
    s.setup_source(src.x, src.w, src.center, src.nq)
    reqs = []
    for seed, N in ((31, 900), (32, 1500)):
        rng = np.random.default_rng(seed)
        y = gmm_sample(rng, N, 3, 4).ravel()
        nu = rng.dirichlet(np.full(N, 5.0))
        reqs.append((y, nu / nu.sum()))
    eps, tol = 1e-2, 1e-6
    res = {}
    for tag, kw in (("small", dict(pass_targets=8, grad_qpoints=32)), ("default", {})):
        for prec in ("fp64", "fp32_lse"):
            with sd.SotServer(s, precision=prec, max_slots=4, hbm_budget_bytes=1 << 30,
                              **kw) as srv:
                ids = [srv.submit(y, nu, eps, tol, "componentwise", 500,
                                  preconditioner="diagonal") for y, nu in reqs]
                out = srv.run(ids)
                res[tag, prec] = [out[i] for i in ids]
    for k, (y, nu) in enumerate(reqs):
        for prec in ("fp64", "fp32_lse"):
            (r_s, psi_s), (r_d, psi_d) = res["small", prec][k], res["default", prec][k]
            _, go, _ = O.evaluate_R(src, O.Target(y, nu), O.params(epsilon=eps, use_lse=1), DMAX,
                                    psi_s)
            err = float(np.max(np.abs(go) / nu))
            dpsi = np.max(np.abs((psi_s - psi_s.mean()) - (psi_d - psi_d.mean())))
            say(f"request {k} {prec}: small items {r_s.status} it={r_s.iterations} "
                f"oracle relerr {err:.3e}; default it={r_d.iterations}; max|dpsi| {dpsi:.2e}")
            assert r_s.status == r_d.status == "converged"
            assert err <= tol * 1.001
            assert dpsi <= 1e-5 * (psi_d.max() - psi_d.min())
    s.close()
