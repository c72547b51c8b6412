"""GPU tests of the batched solve queue (sdot_server_* in include/sdot.h).

Each served solve runs the algorithm of SotSolver::solve (SotSolver.cc:105-215);
the checks are
  * FP64 serving against the single-solve sdot_solve path and the FP64 oracle
    (reference line search: the same decisions for 20 iterations, potentials
    within 1e-12; noise-aware line search: both converge, iteration counts
    within 3, potentials within 1e-6 after removing the constant, true relative
    marginal error <= tolerance);
  * a solve's result is bitwise the same alone and in a mix of solves that are
    admitted and retired around it, with and without the captured graph;
  * FP32_LSE serving with escalation to FP64 for the last iterations: every
    solve converges and the true (FP64 oracle) relative marginal error at the
    returned potential is within the 1e-6 tolerance (the stopping decision is
    taken on an FP64 evaluation);
  * admission by HBM budget, request validation, truncated candidate sets.
"""
import ctypes as C
import sys

import numpy as np
import pytest

import semidiscreteot_amd as sd
from oracle import oracle as O
from tests.cases import gmm_sample, structured_case

pytestmark = pytest.mark.gpu
DMAX = np.finfo(float).max


def say(msg):
    print(f"[gpu-serve] {msg}", flush=True)
    sys.stdout.flush()


def source(n=8, qorder=2, seed=1):
    src, _ = structured_case(3, n, qorder, 16, seed=seed, gmm_source=True)
    s = sd.SotSolver(3)
    s.setup_source(src.x, src.w, src.center, src.nq)
    return src, s


def request(seed, N, dirichlet=5.0):
    rng = np.random.default_rng(seed)
    y = gmm_sample(rng, N, 3, int(rng.integers(3, 9)))
    nu = rng.dirichlet(np.full(N, dirichlet))
    return y.ravel(), nu / nu.sum()


def true_relerr(src, y, nu, eps, psi, R=DMAX):
    _, go, _ = O.evaluate_R(src, O.Target(y, nu), O.params(epsilon=eps, use_lse=1), R, psi)
    return float(np.max(np.abs(go) / nu))


def test_server_fp64_matches_single_solve():
    """Reference line search: the served solve takes the decisions of
    sdot_solve step for step (potentials agree to rounding after 20
    iterations).  Noise-aware line search: both converge to the tolerance."""
    say("fp64 server vs sdot_solve")
    src, s = source()
    y, nu = request(11, 200)
    eps, tol = 1e-2, 1e-6
    s.setup_target(y, nu)
    base = dict(epsilon=eps, tau=0.0, tolerance=tol, solver_control_type="componentwise",
                use_log_sum_exp_trick=True)
    s.configure(sd.SolverParameters(max_iterations=20, **base))
    psi1 = s.solve()
    info = s.get_last_solve_info()
    with sd.SotServer(s, precision="fp64", max_slots=8, hbm_budget_bytes=1 << 30) as srv:
        rid = srv.submit(y, nu, eps, tol, "componentwise", 20, line_search="reference")
        res, psi2 = srv.run([rid])[rid]
    say(f"20 its: single {info['status']} ev={info['evaluations']}  served {res.status} "
        f"ev={res.evaluations}  max|dpsi| {np.max(np.abs(psi1 - psi2)):.2e}")
    assert (info["iterations"], info["evaluations"]) == (res.iterations, res.evaluations) == (20, 21)
    assert np.max(np.abs(psi1 - psi2)) < 1e-12
    assert abs(res.functional - info["functional"]) <= 1e-13 * abs(info["functional"])

    s.configure(sd.SolverParameters(max_iterations=300, line_search="noise_aware", **base))
    psi1 = s.solve()
    info = s.get_last_solve_info()
    with sd.SotServer(s, precision="fp64", max_slots=8, hbm_budget_bytes=1 << 30) as srv:
        rid = srv.submit(y, nu, eps, tol, "componentwise", 300)
        res, psi2 = srv.run([rid])[rid]
    say(f"to 1e-6: single {info['status']} it={info['iterations']}  served {res.status} "
        f"it={res.iterations} ev={res.evaluations} relerr={res.rel_marginal_error:.3e}")
    assert info["status"] == res.status == "converged"
    assert abs(info["iterations"] - res.iterations) <= 3
    assert np.max(np.abs((psi1 - psi1.mean()) - (psi2 - psi2.mean()))) < 1e-6
    assert res.rel_marginal_error <= tol
    assert true_relerr(src, y, nu, eps, psi2) <= tol * 1.001


def _mix_run(s, reqs, target, use_graph, precision="fp32_lse"):
    """Serve `target` (an index of reqs) after two other solves have started
    and while another is admitted after it; returns {index: (res, psi)}."""
    out = {}
    with sd.SotServer(s, precision=precision, max_slots=4, use_graph=use_graph,
                      hbm_budget_bytes=1 << 30) as srv:
        others = [i for i in range(len(reqs)) if i != target]
        ids = {}
        for i in others[:2]:
            ids[srv.submit(*reqs[i])] = i
        srv.step()
        srv.step()
        ids[srv.submit(*reqs[target])] = target
        srv.step()
        for i in others[2:]:
            ids[srv.submit(*reqs[i])] = i
        for rid, r in srv.run(list(ids)).items():
            out[ids[rid]] = r
    return out


def _alone(s, req, use_graph=True, precision="fp32_lse"):
    with sd.SotServer(s, precision=precision, max_slots=4, use_graph=use_graph,
                      hbm_budget_bytes=1 << 30) as srv:
        rid = srv.submit(*req)
        return srv.run([rid])[rid]


def test_server_alone_vs_mix_bitwise():
    say("alone vs mix")
    src, s = source()
    reqs = []
    for seed, N, eps in [(21, 700, 3e-3), (22, 150, 1e-2), (23, 2500, 1e-2), (24, 64, 3e-3)]:
        y, nu = request(seed, N)
        reqs.append((y, nu, eps, 1e-6, "componentwise", 3000))
    mix = _mix_run(s, reqs, target=0, use_graph=True)
    assert all(r[0].status == "converged" for r in mix.values())
    for i in (0, 2):
        res_a, psi_a = _alone(s, reqs[i])
        res_m, psi_m = mix[i]
        say(f"req {i}: alone {res_a.status} it={res_a.iterations} ev={res_a.evaluations}; "
            f"mix {res_m.status} it={res_m.iterations} ev={res_m.evaluations}")
        assert res_a.iterations == res_m.iterations and res_a.evaluations == res_m.evaluations
        assert res_a.functional == res_m.functional
        assert np.array_equal(psi_a, psi_m)
    # direct launches give the same bits as the captured graph
    res_d, psi_d = _alone(s, reqs[0], use_graph=False)
    assert np.array_equal(psi_d, mix[0][1])


def test_server_fp32_true_marginal_error():
    say("fp32 served solves against the FP64 oracle")
    src, s = source()
    reqs = []
    for seed, N, eps in [(31, 300, 1e-2), (32, 1200, 3e-3), (33, 500, 1e-2)]:
        y, nu = request(seed, N)
        reqs.append((y, nu, eps))
    with sd.SotServer(s, precision="fp32_lse", max_slots=8, hbm_budget_bytes=1 << 30) as srv:
        ids = [srv.submit(y, nu, eps, 1e-6, "componentwise", 500) for y, nu, eps in reqs]
        out = srv.run(ids)
    for rid, (y, nu, eps) in zip(ids, reqs):
        res, psi = out[rid]
        err = true_relerr(src, y, nu, eps, psi)
        say(f"N={nu.size} eps={eps}: {res.status} it={res.iterations} "
            f"fp32 relerr {res.rel_marginal_error:.3e}  fp64 relerr {err:.3e}")
        assert res.status == "converged"
        assert err <= 1e-6 * 1.001
        assert abs(err - res.rel_marginal_error) <= 1e-9


def test_server_admission_by_budget():
    say("admission by HBM budget")
    src, s = source()
    reqs = [request(40 + i, 400) for i in range(4)]
    with sd.SotServer(s, precision="fp32_lse", max_slots=16) as probe:
        per = probe.request_bytes(400)
    budget = max(1 << 20, int(per * 2.5))
    with sd.SotServer(s, precision="fp32_lse", max_slots=16, hbm_budget_bytes=budget) as srv:
        with pytest.raises(sd.SdotError):
            y, nu = request(49, 40000)
            srv.submit(y, nu, 1e-2)
        ids = [srv.submit(y, nu, 1e-2, 1e-6, "componentwise", 200) for y, nu in reqs]
        peak = 0
        pending = set(ids)
        results = {}
        while pending:
            na, nq = srv.step()
            peak = max(peak, na)
            for rid in list(pending):
                r = srv.poll(rid)
                if r is not None:
                    results[rid] = r
                    pending.discard(rid)
        say(f"per-request bytes {per}, budget {budget}, peak active {peak}")
        assert peak == 2
        assert all(r[0].status == "converged" for r in results.values())
        # admission follows submission order
        admits = [results[i][0].t_admit for i in ids]
        assert admits == sorted(admits)


def test_server_truncated_radius_and_errors():
    say("truncated radius")
    src, s = source()
    y, nu = request(51, 600)
    eps, R = 1e-2, 0.7
    with sd.SotServer(s, precision="fp64", max_slots=4, hbm_budget_bytes=1 << 30) as srv:
        rid = srv.submit(y, nu, eps, 1e-6, "componentwise", 300, radius=R)
        res, psi = srv.run([rid])[rid]
        Fo, go, _ = O.evaluate_R(src, O.Target(y, nu), O.params(epsilon=eps, use_lse=1), R, psi)
        scale = max(abs(Fo), abs(float(np.dot(psi, nu))))
        say(f"{res.status} it={res.iterations} F={res.functional!r} oracle F={Fo!r}")
        assert res.status == "converged"
        assert abs(res.functional - Fo) <= 1e-10 * scale
        assert abs(res.rel_marginal_error - np.max(np.abs(go) / nu)) <= 1e-9
        with pytest.raises(sd.SdotError):
            srv.submit(y, nu, -1.0)
        with pytest.raises(sd.SdotError):
            srv.submit(y, -nu, 1e-2)
        done = C.c_int32(7)
        assert srv.lib.sdot_server_poll(srv._h, 12345, C.byref(done), None, None) != 0
        # a finished request is forgotten once polled
        assert srv.lib.sdot_server_poll(srv._h, rid, C.byref(done), None, None) == 0
        assert done.value == 0
    # the source of a context lent to a server cannot change meanwhile
    with sd.SotServer(s, precision="fp64") as srv:
        with pytest.raises(sd.SdotError):
            s.setup_source(src.x, src.w, src.center, src.nq)
    s.setup_source(src.x, src.w, src.center, src.nq)


def test_server_autotune():
    say("autotune")
    src, s = source()
    with sd.SotServer(s, precision="fp32_lse", max_slots=8, hbm_budget_bytes=4 << 30) as srv:
        pt, gq = srv.autotune()
        say(f"pass_targets={pt} grad_qpoints={gq}")
        assert pt in (1024, 2048, 4096, 8192) and gq in (2048, 4096, 8192, 16384)
        y, nu = request(61, 300)
        rid = srv.submit(y, nu, 1e-2, 1e-6, "componentwise", 300)
        res, psi = srv.run([rid])[rid]
        assert res.status == "converged"


def test_server_epsilon_scaling_matches_run_sot_loop():
    """Epsilon scaling (SemiDiscreteOT.cc:877-918): the served request runs the
    same stages as SotSolver.solve_epsilon_scaling; both converge at the final
    epsilon with the tolerance met in FP64."""
    say("epsilon scaling")
    src, s = source()
    y, nu = request(71, 800)
    eps, tol, steps, factor = 3e-3, 1e-6, 4, 2.0
    s.setup_target(y, nu)
    prm = sd.SolverParameters(epsilon=eps, tau=0.0, tolerance=tol,
                              solver_control_type="componentwise", max_iterations=3000,
                              use_log_sum_exp_trick=True, line_search="noise_aware",
                              use_epsilon_scaling=True, epsilon_scaling_steps=steps,
                              epsilon_scaling_factor=factor)
    psi1, infos = s.solve_epsilon_scaling(None, prm)
    assert [i["epsilon"] for i in infos] == [eps * 8, eps * 4, eps * 2, eps]
    it1 = sum(i["iterations"] for i in infos)
    with sd.SotServer(s, precision="fp64", max_slots=4, hbm_budget_bytes=1 << 30) as srv:
        rid = srv.submit(y, nu, eps, tol, "componentwise", 3000, epsilon_scaling_steps=steps,
                         epsilon_scaling_factor=factor)
        res, psi2 = srv.run([rid])[rid]
    say(f"run_sot loop: {[i['status'] for i in infos]} {it1} its; served {res.status} "
        f"{res.iterations} its, relerr {res.rel_marginal_error:.3e}")
    assert infos[-1]["status"] == res.status == "converged"
    assert abs(it1 - res.iterations) <= 3 * steps
    assert np.max(np.abs((psi1 - psi1.mean()) - (psi2 - psi2.mean()))) < 1e-6
    assert true_relerr(src, y, nu, eps, psi2) <= tol * 1.001


def test_diagonal_preconditioner_single_and_served():
    """Initial inverse Hessian diag(eps / nu_j): the served FP64 solve takes the
    decisions of sdot_solve step for step (20 iterations, potentials to
    rounding); solved to 1e-6 (noise-aware search) it converges in fewer
    iterations than with the identity, both to the tolerance on the FP64
    oracle's marginals and to the same potential; an FP32 served solve reaches
    the tolerance on the oracle's marginals too."""
    say("diagonal preconditioner")
    src, s = source()
    y, nu = request(12, 400, dirichlet=2.0)
    eps, tol = 3e-3, 1e-6
    s.setup_target(y, nu)
    base = dict(epsilon=eps, tau=0.0, tolerance=tol, solver_control_type="componentwise",
                use_log_sum_exp_trick=True, preconditioner="diagonal")
    s.configure(sd.SolverParameters(max_iterations=20, **base))
    psi1 = s.solve()
    info = s.get_last_solve_info()
    with sd.SotServer(s, precision="fp64", max_slots=8, hbm_budget_bytes=1 << 30) as srv:
        rid = srv.submit(y, nu, eps, tol, "componentwise", 20, line_search="reference",
                         preconditioner="diagonal")
        res, psi2 = srv.run([rid])[rid]
    assert (info["iterations"], info["evaluations"]) == (res.iterations, res.evaluations)
    assert np.max(np.abs(psi1 - psi2)) < 1e-12

    its, sol = {}, {}
    for pc in ("none", "diagonal"):
        s.configure(sd.SolverParameters(max_iterations=2000, line_search="noise_aware",
                                        **{**base, "preconditioner": pc}))
        sol[pc] = s.solve()
        inf = s.get_last_solve_info()
        its[pc] = inf["iterations"]
        assert inf["status"] == "converged"
        assert true_relerr(src, y, nu, eps, sol[pc]) <= tol * 1.001
    a, b = sol["none"], sol["diagonal"]
    dpsi = np.max(np.abs((a - a.mean()) - (b - b.mean())))
    say(f"iterations: identity {its['none']}, diagonal {its['diagonal']}; |dpsi| {dpsi:.2e}")
    assert its["diagonal"] < its["none"]
    assert dpsi < 1e-5

    with sd.SotServer(s, precision="fp32_lse", max_slots=8, hbm_budget_bytes=1 << 30) as srv:
        rid = srv.submit(y, nu, eps, tol, "componentwise", 2000, preconditioner="diagonal")
        res, psi3 = srv.run([rid])[rid]
    err = true_relerr(src, y, nu, eps, psi3)
    say(f"fp32 served diagonal: {res.status} it={res.iterations} fp64 relerr {err:.3e}")
    assert res.status == "converged" and err <= tol * 1.001
