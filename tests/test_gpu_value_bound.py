"""The value bound of the pair sums (`prune_bits`): bins of candidate pairs whose
terms all lie below 2^-bits of their q-point's largest weighted term are
skipped.  Its guarantee, for every q-point and target,

    |S_q(bounded) - S_q| <= N 2^-bits S_q      (the q-point's sum)
    |m_j(bounded) - m_j| <= 2^-bits + N 2^-bits m_j   (a target's mass, sum w = 1)

is checked against the FP64 oracle (every candidate pair) with a bound loose
enough to change the result measurably (20 bits at eps = 2e-2), in FP64 so
that nothing but the bound separates the two; FP64 at 100 bits must stay within the parity bar
(1e-10 relative); FP32_LSE with its default 50 bits within the FP32 bars
(1e-6 of max(nu_j, m_j) on g, 1e-8 relative on F), dense and truncated, across
a seeded sequence of device evaluations and through the re-sweep of an item
whose seed lies too high."""
import os
import sys

import numpy as np
import pytest

import semidiscreteot_amd as sd
from semidiscreteot_amd import devmem
from oracle import oracle as O
from tests.cases import structured_case

pytestmark = pytest.mark.gpu

DMAX = float(np.finfo(float).max)
EPS = 1e-3


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def case(N=3000, seed=31):
    return structured_case(3, 12, 2, N, seed=seed, gmm_source=True, gmm_targets=True,
                           dirichlet=5)


def solver(src, tgt, precision, bits):
    s = sd.SotSolver(3)
    s.setup_source(src.x, src.w, src.center, src.nq)
    s.setup_target(tgt.y, tgt.nu)
    s.configure(sd.SolverParameters(epsilon=EPS, use_log_sum_exp_trick=True, precision=precision,
                                    prune_bits=bits, verbose_output=False))
    return s


def oracle(src, tgt, R, psi):
    return O.evaluate_R(src, tgt, O.params(epsilon=EPS, use_lse=1), R, psi)


def device_eval(s, R, psi):
    d_psi = devmem.DeviceArray.from_host(psi)
    d_g = devmem.DeviceArray(psi.size)
    d_fc = devmem.DeviceArray(4)
    s.eval_device(R, d_psi.address, d_g.address, d_fc.address)
    s.synchronize()
    fc = d_fc.download()
    return fc[0], d_g.download(), fc[1]


@pytest.mark.parametrize("R", [0.9, DMAX], ids=["truncated", "dense"])
def test_bound_holds_and_is_active_fp64(R):
    """20 bits at eps = 2e-2, targets spread over [-0.5, 1.5]^3 around the unit
    cube source: the skipped bins lie just past the radius sqrt(2 * 20 ln2 eps)
    = 0.74 where a target's terms fall below 2^-20, so the skipped terms are
    large enough to move the FP64 result."""
    eps = 2e-2
    src, _ = case(N=1000, seed=31)
    rng = np.random.default_rng(1)
    y = rng.uniform(-0.5, 1.5, size=(1000, 3))
    nu = rng.dirichlet(np.full(1000, 5.0))
    tgt = O.Target(y.ravel(), nu / nu.sum())
    psi = 0.01 * rng.standard_normal(tgt.n)
    Fo, go, _ = O.evaluate_R(src, tgt, O.params(epsilon=eps, use_lse=1), R, psi)
    mo = go + tgt.nu
    bits = 20
    s = sd.SotSolver(3)
    s.setup_source(src.x, src.w, src.center, src.nq)
    s.setup_target(tgt.y, tgt.nu)
    s.configure(sd.SolverParameters(epsilon=eps, use_log_sum_exp_trick=True, precision="fp64",
                                    prune_bits=bits, verbose_output=False))
    say(f"FP64, {bits}-bit value bound, eps {eps:g}, R={R:g}")
    F, g, _ = s.evaluate_functional_R(R, psi)
    s.close()
    dm = np.abs(g - go)
    tol_m = 2.0 ** -bits + 1.01 * tgt.n * 2.0 ** -bits * mo + 1e-12 * np.maximum(mo, tgt.nu)
    dF = abs(F - Fo)
    tol_F = 1.01 * eps * tgt.n * 2.0 ** -bits + 1e-12 * abs(Fo)
    say(f"max |dm| {dm.max():.3e} (bound {tol_m.max():.3e}), |dF| {dF:.3e} (bound {tol_F:.3e})")
    assert np.all(dm <= tol_m)
    assert dF <= tol_F
    # the bound skipped pairs: the result differs from every-pair sums by more
    # than their rounding
    assert dm.max() > 1e-12 * mo.max() and dF > 1e-13 * abs(Fo)


@pytest.mark.parametrize("R", [0.45, DMAX], ids=["truncated", "dense"])
def test_fp64_at_100_bits_keeps_parity(R):
    src, tgt = case(seed=32)
    rng = np.random.default_rng(2)
    psi = 0.01 * rng.standard_normal(tgt.n)
    Fo, go, Co = oracle(src, tgt, R, psi)
    s = solver(src, tgt, "fp64", 100)
    F, g, C = s.evaluate_functional_R(R, psi)
    s.close()
    relF = abs(F - Fo) / abs(Fo)
    relg = np.max(np.abs(g - go)) / np.max(np.abs(go + tgt.nu))
    relC = abs(C - Co) / abs(Co)
    say(f"FP64 100 bits R={R:g}: F {relF:.2e}, g {relg:.2e}, C {relC:.2e}")
    assert relF <= 1e-10 and relg <= 1e-10 and relC <= 1e-10


@pytest.mark.parametrize("R", [0.45, DMAX], ids=["truncated", "dense"])
def test_fp32_bounded_sequence_vs_oracle(R):
    """Seeded FP32 device evaluations with the 50-bit bound along a sequence of
    potentials (small steps: seeded items; one jump: unseeded)."""
    src, tgt = case(seed=33)
    rng = np.random.default_rng(3)
    psi = 0.01 * rng.standard_normal(tgt.n)
    s = solver(src, tgt, "fp32_lse", 50)
    for step in (0.0, 1e-4, 1e-4, 2e-2, 1e-5):
        psi = psi + step * rng.standard_normal(tgt.n)
        F, g, _ = device_eval(s, R, psi)
        Fo, go, _ = oracle(src, tgt, R, psi)
        relm = np.max(np.abs(g - go) / np.maximum(tgt.nu, go + tgt.nu))
        relF = abs(F - Fo) / abs(Fo)
        say(f"FP32 50 bits R={R:g} step {step:g}: g {relm:.2e}, F {relF:.2e}")
        assert relm <= 1e-6 and relF <= 1e-8
    s.close()


def test_fp32_bounded_reswept_item_equals_cold():
    """An item whose seed lies 40 log2 units too high (test hook) is swept again
    after an exact maximum over every candidate; with the value bound the
    result still equals a cold evaluation bitwise."""
    src, tgt = case(seed=34)
    R = DMAX
    rng = np.random.default_rng(4)
    psi_a = 0.01 * rng.standard_normal(tgt.n)
    psi_b = psi_a + 1e-4 * rng.standard_normal(tgt.n)
    os.environ["SDOT_TEST_SEED_BIAS"] = "40"
    try:
        s = solver(src, tgt, "fp32_lse", 50)
    finally:
        del os.environ["SDOT_TEST_SEED_BIAS"]
    device_eval(s, R, psi_a)
    F1, g1, C1 = device_eval(s, R, psi_b)
    cold = solver(src, tgt, "fp32_lse", 50)
    F2, g2, C2 = device_eval(cold, R, psi_b)
    say(f"re-swept vs cold: F {F1!r} {F2!r}")
    assert F1 == F2 and C1 == C2 and np.array_equal(g1, g2)
    s.close()
    cold.close()


def test_default_bound_where_it_pays():
    """The FP32_LSE default bounds large evaluations at small epsilon only; its
    result matches the unbounded FP32 evaluation to FP32 accuracy and FP64
    stays unbounded (bitwise equal to prune_bits = 0)."""
    src, tgt = structured_case(3, 16, 2, 140000, seed=35, gmm_source=True, gmm_targets=True,
                               dirichlet=5)
    rng = np.random.default_rng(5)
    psi = 0.005 * rng.standard_normal(tgt.n)
    out = {}
    for prec, bits in (("fp32_lse", -1), ("fp32_lse", 0), ("fp64", -1), ("fp64", 0)):
        s = solver(src, tgt, prec, bits)
        out[(prec, bits)] = s.evaluate_functional_R(DMAX, psi)
        s.close()
    F1, g1, _ = out[("fp32_lse", -1)]
    F0, g0, _ = out[("fp32_lse", 0)]
    m0 = g0 + tgt.nu
    relm = np.max(np.abs(g1 - g0) / np.maximum(tgt.nu, m0))
    say(f"default vs unbounded FP32: g {relm:.2e}, F {abs(F1 - F0) / abs(F0):.2e}")
    assert relm <= 1e-7 and abs(F1 - F0) <= 1e-9 * abs(F0)
    assert not np.array_equal(g1, g0)  # the default did bound this evaluation
    assert out[("fp64", -1)][0] == out[("fp64", 0)][0]
    assert np.array_equal(out[("fp64", -1)][1], out[("fp64", 0)][1])


def test_served_bounded_request_mix_vs_alone_and_oracle():
    """A served FP32 request that the default bounds (20,000 targets at eps 1e-3
    against the 262,144 q-points of the 32^3 source: 5.2e9 pairs) beside one
    it does not bound (3,000 targets): each result is bitwise the same alone
    and in the mix, and within 1e-6 relative marginal error of the FP64
    oracle at its potential."""
    from semidiscreteot_amd import workloads
    wl = workloads.gmm_3d()
    s = sd.SotSolver(3)
    s.setup_source(wl.source.x, wl.source.w, wl.source.centers, wl.source.nq)
    rng = np.random.default_rng(36)
    reqs = []
    for n in (20000, 3000):
        y = workloads.gmm_points(rng, n, 3, 5)
        nu = rng.dirichlet(np.full(n, 2.0))
        reqs.append((y.ravel(), nu / nu.sum()))
    kw = dict(control="componentwise", max_iterations=1000, preconditioner="diagonal")

    def serve(rs):
        with sd.SotServer(s, precision="fp32_lse", max_slots=8, hbm_budget_bytes=8 << 30) as srv:
            ids = [srv.submit(y, nu, EPS, 1e-6, **kw) for y, nu in rs]
            out = srv.run(ids)
            return [out[i] for i in ids]

    mix = serve(reqs)
    alone = [serve([r])[0] for r in reqs]
    src = O.Source(3, wl.source.nq, wl.source.x, wl.source.w, np.ones_like(wl.source.w),
                   wl.source.centers)
    for (y, nu), (rm, pm), (ra, pa) in zip(reqs, mix, alone):
        _, go, _ = O.evaluate_R(src, O.Target(y, nu), O.params(epsilon=EPS, use_lse=1), DMAX, pm)
        err = float(np.max(np.abs(go) / nu))
        say(f"served N {nu.size}: {rm.status} after {rm.iterations} iterations, oracle relerr "
            f"{err:.3e}, bitwise alone/mix {np.array_equal(pm, pa)}")
        assert rm.status == ra.status == "converged"
        assert np.array_equal(pm, pa) and rm.iterations == ra.iterations
        assert err <= 1e-6
    s.close()


@pytest.mark.parametrize("precision, bits", [("fp64", 100.0), ("fp32_lse", 50.0)])
@pytest.mark.parametrize("R", [0.45, DMAX], ids=["truncated", "dense"])
def test_bounded_barycenters_vs_oracle(precision, bits, R):
    """The Lloyd step's weighted barycenters run the bounded gradient sweep with
    the first moments: FP64 at 100 bits within 1e-10 of the oracle, FP32 with
    the 50-bit bound within the FP32 barycenter bar (2e-6)."""
    src, tgt = case(N=1500, seed=37)
    rng = np.random.default_rng(7)
    psi = 0.01 * rng.standard_normal(tgt.n)
    s = solver(src, tgt, precision, bits)
    b = s.compute_weighted_barycenters_R(R, psi)
    s.close()
    bo = O.barycenters_R(src, tgt, O.params(epsilon=EPS, use_lse=1), R, psi)
    err = float(np.max(np.abs(b - bo)))
    say(f"barycenters {precision} {bits:g} bits R={R:g}: max |b - b_oracle| {err:.3e}")
    if precision == "fp64":
        assert err <= 1e-10 * max(1.0, float(np.max(np.abs(bo))))
    else:
        assert err <= 2e-6
