"""GPU point queries through the C ABI against the FP64 oracle and the
50-digit golden values: conditional densities
(SotSolver::get_potential_conditioned_density, SotSolver.cc:973-1046) and the
modal / barycentric transport maps (OptimalTransportPlan.cc:99-189, 218-310).

Tolerances: conditional densities and barycentric maps
max |v - v_ref| <= 1e-12 * max(1, max |v_ref|) with the same zero pattern
(candidate sets are exact); modal maps and nearest-target fallbacks are
indices, so the mapped points are equal bit for bit."""
import json
import os
import sys

import numpy as np
import pytest

import semidiscreteot_amd as sd
from oracle import oracle as O
from tests.cases import structured_case

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
POINTS = json.load(open(os.path.join(HERE, "golden", "points_cases.json")))
DMAX = np.finfo(float).max


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def solver_for(dim, y, nu, eps):
    """A context whose targets are (y, nu); its source is a small box mesh
    (the point queries read only the targets and the configured epsilon)."""
    src, _ = structured_case(dim, 4, 2, 1, seed=0)
    s = sd.SotSolver(dim)
    s.setup_source(src.x, src.w, src.center, src.nq)
    s.setup_target(np.ravel(y), np.ravel(nu))
    s.configure(sd.SolverParameters(epsilon=eps, use_log_sum_exp_trick=True))
    return s


def close(v, ref, tol=1e-12):
    assert np.array_equal(v == 0.0, ref == 0.0)
    assert np.max(np.abs(v - ref)) <= tol * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("case", POINTS, ids=[c["name"] for c in POINTS])
def test_points_golden(case):
    say(f"points golden {case['name']}")
    dim = case["dim"]
    y = np.asarray(case["y"], dtype=float)
    nu = np.asarray(case["nu"], dtype=float)
    psi = np.asarray(case["psi"], dtype=float)
    pts = np.asarray(case["points"], dtype=float).reshape(-1, dim)
    s = solver_for(dim, y, nu, case["eps"])
    cond = s.get_potential_conditioned_density(pts, case["rho"], psi, case["indices"],
                                               R=case["R"])
    ref = np.array([[float(v) for v in r] for r in case["cond"]])
    close(cond, ref, 1e-13)
    s.close()
    plan = sd.OptimalTransportPlan(dim, "barycentric")
    plan.set_source_measure(pts, case["rho"])
    plan.set_target_measure(y.reshape(-1, dim), nu)
    plan.set_potential(psi, case["eps"])
    plan.set_truncation_radius(case["radius"])
    m = plan.compute_map()
    rb = np.array([[float(v) for v in r] for r in case["bary_map"]])
    assert np.max(np.abs(m - rb)) <= 1e-14
    assert np.array_equal(plan.transport_density, np.asarray(case["rho"], dtype=float))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("truncated", [True, False])
def test_conditional_density_vs_oracle(dim, truncated):
    say(f"conditional density dim={dim} truncated={truncated}")
    rng = np.random.default_rng(10 + dim)
    N, n = 3000, 2500
    y = rng.random((N, dim))
    nu = rng.dirichlet(np.full(N, 3.0))
    psi = 0.01 * rng.standard_normal(N)
    pts = rng.random((n, dim))
    rho = 1.0 + rng.random(n)
    idx = np.array([0, 17, N // 2, N - 1, -3, N + 5], dtype=np.int64)
    eps = 4e-3
    R = 0.2 if truncated else DMAX
    s = solver_for(dim, y, nu, eps)
    c = s.get_potential_conditioned_density(pts, rho, psi, idx, R=R)
    tgt = O.Target(y.ravel(), nu)
    co = O.conditional_density(tgt, dim, psi, eps, R, pts, rho, idx)
    close(c, co)
    assert np.all(c[4:] == 0.0)  # indices outside the target set
    if not truncated:
        # partition of unity over every target: sum_i nu_i c_i(p) = rho(p)
        allc = s.get_potential_conditioned_density(pts[:200], rho[:200], psi, np.arange(N), R=R)
        assert np.max(np.abs((nu[:, None] * allc).sum(0) - rho[:200])) <= 1e-13
    s.close()


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("strategy", ["modal", "barycentric"])
@pytest.mark.parametrize("radius", [0.0, 0.05])
def test_transport_map_vs_oracle(dim, strategy, radius):
    """radius 0: every target; 0.05: truncated, with query points far outside
    the targets' box that have no candidate and take the nearest target."""
    say(f"transport map dim={dim} {strategy} radius={radius}")
    rng = np.random.default_rng(20 + dim)
    N, n = 4000, 3000
    y = rng.random((N, dim))
    nu = rng.dirichlet(np.full(N, 2.0))
    nu[5] = 0.0  # -inf log term in the modal score
    eps = 3e-3
    psi = 0.02 * rng.standard_normal(N)
    pts = rng.random((n, dim))
    pts[:40] = 1.5 + rng.random((40, dim))   # no target within the radius
    pts[40:50] = y[100:110]                  # exactly on targets
    plan = sd.OptimalTransportPlan(dim, strategy)
    plan.set_source_measure(pts, np.ones(n))
    plan.set_target_measure(y, nu)
    plan.set_potential(psi, eps)
    plan.set_truncation_radius(radius)
    m = plan.compute_map()
    mo = O.transport_map(0 if strategy == "modal" else 1, O.Target(y.ravel(), nu), dim, psi, eps,
                         radius, pts)
    if radius > 0:
        assert np.array_equal(m[:40], mo[:40])  # nearest-target fallback
    if strategy == "modal":
        assert np.array_equal(m, mo)
    else:
        assert np.max(np.abs(m - mo)) <= 1e-12


def test_point_query_errors():
    say("point query errors")
    plan = sd.OptimalTransportPlan(2, "modal")
    with pytest.raises(RuntimeError):
        plan.compute_map()
    with pytest.raises(RuntimeError):
        sd.OptimalTransportPlan(2, "nearest")
    assert sd.OptimalTransportPlan.get_available_strategies() == ["modal", "barycentric"]
