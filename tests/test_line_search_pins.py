"""Hand-derived traces of the L-BFGS driver and its line search, run on both
FP64 restatements of deal.II's SolverBFGS / LineMinimization::line_search:
the oracle's (oracle/sdot_oracle.c, orc_bfgs_generic) and the library's
(sdot_bfgs.hpp lbfgs_solve, the host driver of sdot_solve; through
tests/ls_script_check.cpp).

Each case is a scripted objective: the k-th evaluation returns the k-th
(f, g) of the script whatever point it is asked at, so the trial points the
search asks for are its decisions.  The expected points are worked out by
hand from the algorithm's text (core/line_minimization.h:498-749, poly_fit
:423-455; first step core/wasserstein_solver_bfgs.h:335-341; two-loop
recursion :427-448) and written down as numbers; where a number comes from
interpolation it is checked against an independent construction (Hermite
cubic and parabola through the bracket data, solved with numpy).  The cases
are chosen so that changing any one constant of the text -- growth 9,
safeguards 0.1 / 0.5, first-step factor 2.02, Wolfe eta 0.9 and mu 1e-3, the
20-evaluation cap -- or a slip in the cubic, quadratic or two-loop formulas
moves a trial point.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STATUS = {0: "converged", 1: "max_iterations", 2: "linesearch_fail", 3: "nonfinite"}


def run_oracle(script, x0, max_it, tol=1e-14):
    xs = []
    k = [0]

    def fun(x):
        xs.append(x.copy())
        f, g = script[min(k[0], len(script) - 1)]
        k[0] += 1
        return f, np.asarray(g, dtype=float)

    _, info = O.bfgs_generic(fun, np.asarray(x0, dtype=float), max_iterations=max_it, tol=tol)
    return np.array(xs), info["status"], info["iterations"], info["evaluations"]


@pytest.fixture(scope="module")
def product_exe(tmp_path_factory):
    exe = tmp_path_factory.mktemp("ls") / "ls"
    subprocess.run(["g++", "-std=c++17", "-O2", os.path.join(ROOT, "tests", "ls_script_check.cpp"),
                    "-o", str(exe)], check=True)
    return str(exe)


def run_product(exe, script, x0, max_it, tol=1e-14):
    n = len(x0)
    lines = [f"{n} {max_it} {tol!r}", " ".join(repr(float(v)) for v in x0), str(len(script))]
    for f, g in script:
        lines.append(" ".join(repr(float(v)) for v in [f, *np.atleast_1d(g)]))
    r = subprocess.run([exe], input="\n".join(lines) + "\n", capture_output=True, text=True,
                       check=True)
    rows = r.stdout.strip().splitlines()
    st, it, ev = (int(v) for v in rows[-1].split())
    xs = np.array([[float(v) for v in row.split()] for row in rows[:-1]])
    return xs, STATUS[st], it, ev


def both(exe, script, x0, max_it):
    return [run_oracle(script, x0, max_it), run_product(exe, script, x0, max_it)]


def hermite_min(x1, f1, g1, x2, f2, g2):
    """Local minimiser of the cubic with the given values and slopes at x1, x2
    (independent of the closed form the line search uses)."""
    A = np.array([[1, x1, x1 ** 2, x1 ** 3], [0, 1, 2 * x1, 3 * x1 ** 2],
                  [1, x2, x2 ** 2, x2 ** 3], [0, 1, 2 * x2, 3 * x2 ** 2]], dtype=float)
    c = np.linalg.solve(A, [f1, g1, f2, g2])
    roots = np.roots([3 * c[3], 2 * c[2], c[1]])
    mins = [r.real for r in roots if abs(r.imag) < 1e-12 and 2 * c[2] + 6 * c[3] * r.real > 0]
    return mins[0] if mins else None


def parabola_vertex(x1, f1, g1, x2, f2):
    """Vertex of the parabola through (x1, f1) with slope g1 and through (x2, f2)."""
    c = (f2 - f1 - g1 * (x2 - x1)) / (x2 - x1) ** 2
    return x1 - g1 / (2 * c)


# 1-D scripts start at x0 = 0 with g0 = -1, so the first direction is p = +1
# and a trial step a is the point x = a.

def test_accepts_first_trial_when_both_wolfe_conditions_hold(product_exe):
    # a = 1: f = -0.5 <= 0 + 1 * 1e-3 * (-1); |g| = 0.895 <= 0.9 * 1 -> accept
    for xs, st, it, ev in both(product_exe, [(0.0, -1.0), (-0.5, -0.895)], [0.0], 1):
        assert xs[:, 0].tolist() == [0.0, 1.0]
        assert (st, it, ev) == ("max_iterations", 1, 2)


def test_curvature_condition_threshold_eta(product_exe):
    # |g| = 0.905 > 0.9 -> not accepted, slope still negative -> grow to
    # 1 + 9 (1 - 0) = 10, where |g| = 0.5 <= 0.9 accepts
    script = [(0.0, -1.0), (-0.5, -0.905), (-5.0, -0.5)]
    for xs, st, it, ev in both(product_exe, script, [0.0], 1):
        assert xs[:, 0].tolist() == [0.0, 1.0, 10.0]
        assert ev == 3


def test_sufficient_decrease_threshold_mu(product_exe):
    # f(1) = -0.00105 <= -1e-3: accepted at once (g = 0)
    for xs, st, it, ev in both(product_exe, [(0.0, -1.0), (-0.00105, 0.0)], [0.0], 1):
        assert xs[:, 0].tolist() == [0.0, 1.0]
    # f(1) = -0.00095 > -1e-3: bracket [0, 1], safeguarded interval
    # [0 + 0.1, 1 - 0.5] = [0.1, 0.5]; the cubic through (0, 0, -1), (1, -0.00095, 0)
    # has its minimum inside it
    a = hermite_min(0.0, 0.0, -1.0, 1.0, -0.00095, 0.0)
    assert 0.1 < a < 0.5 and abs(a - 0.33397) < 1e-4
    script = [(0.0, -1.0), (-0.00095, 0.0), (-0.5, 0.0)]
    for xs, st, it, ev in both(product_exe, script, [0.0], 1):
        assert xs[:2, 0].tolist() == [0.0, 1.0]
        assert abs(xs[2, 0] - a) < 1e-12


def test_bracketing_growth_then_quadratic_fallback(product_exe):
    # slopes stay at -1 with f decreasing: trials 1, 1 + 9 * 1 = 10,
    # 10 + 9 * (10 - 1) = 91; f(91) = 0 > 0 + 91e-3 * (-1) brackets [10, 91].
    # Safeguards: [10 + 0.1 * 81, 91 - 0.5 * 81] = [18.1, 50.5].  The cubic
    # through (10, -10, -1), (91, 0, 2) has its minimum at 52.22 > 50.5, the
    # parabola through (10, -10, -1), (91, 0) its vertex at 10 + 6561 / 182.
    cub = hermite_min(10.0, -10.0, -1.0, 91.0, 0.0, 2.0)
    assert 50.5 < cub < 53.0
    vx = parabola_vertex(10.0, -10.0, -1.0, 91.0, 0.0)
    assert abs(vx - (10.0 + 6561.0 / 182.0)) < 1e-12
    script = [(0.0, -1.0), (-1.0, -1.0), (-10.0, -1.0), (0.0, 2.0), (-30.0, 0.0)]
    for xs, st, it, ev in both(product_exe, script, [0.0], 1):
        assert xs[:4, 0].tolist() == [0.0, 1.0, 10.0, 91.0]
        assert abs(xs[4, 0] - 46.04945054945055) < 1e-12
        assert ev == 5


def test_bisection_of_the_safeguarded_interval(product_exe):
    # the same bracket with f(91) = 1e6, g(91) = 0: the cubic's minimum (10.0011)
    # and the parabola's vertex (10.0031) both fall below 18.1, so the trial is
    # the midpoint of [18.1, 50.5]: 34.3
    cub = hermite_min(10.0, -10.0, -1.0, 91.0, 1e6, 0.0)
    vx = parabola_vertex(10.0, -10.0, -1.0, 91.0, 1e6)
    assert cub < 18.1 and vx < 18.1
    script = [(0.0, -1.0), (-1.0, -1.0), (-10.0, -1.0), (1e6, 0.0), (-30.0, 0.0)]
    for xs, st, it, ev in both(product_exe, script, [0.0], 1):
        assert xs[:4, 0].tolist() == [0.0, 1.0, 10.0, 91.0]
        assert abs(xs[4, 0] - 34.3) < 1e-12


def test_cubic_interpolation_is_exact_on_a_cubic(product_exe):
    # f(x) = x^3 / 3 - 16 x from x0 = 0: g0 = -16, p = 16, trial x = 16 with
    # f = 4096/3 - 256 > 0 brackets [0, 1] in a; the cubic fit reproduces
    # phi(a) = 4096/3 a^3 - 256 a, minimum a = 1/4, x = 4, g(4) = 0
    f = lambda x: x ** 3 / 3.0 - 16.0 * x
    script = [(f(0.0), -16.0), (f(16.0), 240.0), (f(4.0), 0.0)]
    for xs, st, it, ev in both(product_exe, script, [0.0], 5):
        assert xs[:2, 0].tolist() == [0.0, 16.0]
        assert abs(xs[2, 0] - 4.0) < 1e-12
        assert (st, ev) == ("converged", 3)


def test_first_step_rule_of_later_iterations(product_exe):
    # iteration 1 accepts a = 1 (f = -0.1, g = -0.5).  History s = 1, y = 0.5:
    # the two-loop direction is -(s / y) g = +1, g0 = -0.5; the first trial is
    # min(1, -2.02 * (0 - (-0.1)) / (-0.5)) = 0.404, i.e. x = 1.404
    script = [(0.0, -1.0), (-0.1, -0.5), (-0.2, 0.0)]
    for xs, st, it, ev in both(product_exe, script, [0.0], 5):
        assert xs[:2, 0].tolist() == [0.0, 1.0]
        assert abs(xs[2, 0] - 1.404) < 1e-12
        assert (st, it, ev) == ("converged", 2, 3)


def test_evaluation_cap_fails_the_line_search(product_exe):
    # f never decreases: the search brackets [0, 1] and sections until its 20
    # evaluations are spent, then the solve stops with a line-search failure
    script = [(0.0, -1.0), (1.0, 1.0)]
    for xs, st, it, ev in both(product_exe, script, [0.0], 100):
        assert (st, it, ev) == ("linesearch_fail", 1, 21)
        # every sectioning trial lies in the safeguarded part of [0, previous]
        a = xs[1:, 0]
        assert a[0] == 1.0 and np.all(a[1:] > 0.0) and np.all(np.diff(a) < 0)


def test_two_loop_direction_equals_explicit_bfgs_matrix(product_exe):
    # 2-D.  Iteration 1: g0 = (-1, 0), p = (1, 0), accepted at a = 1 with
    # f = -0.6, g1 = (-0.5, -0.5).  s0 = (1, 0), y0 = (0.5, -0.5), rho0 = 2:
    # H1 = V0' V0 + rho0 s0 s0' = [[3, 1], [1, 1]] (H1 y0 = s0), p1 = -H1 g1 = (2, 1).
    # Iteration 2 starts at min(1, 2.02 * 0.6 / 1.5) = 0.808: x = (2.616, 0.808),
    # accepted with f = -1, g2 = (0.1, -0.3); iteration 3 uses H2 from the
    # explicit update with (s1, y1).
    H1 = np.array([[3.0, 1.0], [1.0, 1.0]])
    s0, y0 = np.array([1.0, 0.0]), np.array([0.5, -0.5])
    assert np.allclose(H1 @ y0, s0)

    def update(H, s, y):
        r = 1.0 / (s @ y)
        V = np.eye(2) - r * np.outer(y, s)
        return V.T @ H @ V + r * np.outer(s, s)

    assert np.allclose(update(np.eye(2), s0, y0), H1)
    g1 = np.array([-0.5, -0.5])
    p1 = -H1 @ g1
    assert np.allclose(p1, [2.0, 1.0])
    x2 = np.array([1.0, 0.0]) + 0.808 * p1
    g2 = np.array([0.1, -0.3])
    s1, y1 = x2 - np.array([1.0, 0.0]), g2 - g1
    H2 = update(H1, s1, y1)
    assert np.allclose(H2 @ y1, s1)
    p2 = -H2 @ g2
    a2 = min(1.0, -2.02 * (-0.6 - (-1.0)) / (g2 @ p2))
    x3 = x2 + a2 * p2
    script = [(0.0, [-1.0, 0.0]), (-0.6, [-0.5, -0.5]), (-1.0, [0.1, -0.3]), (-1.2, [0.0, 0.0])]
    for xs, st, it, ev in both(product_exe, script, [0.0, 0.0], 5):
        assert np.allclose(xs[0], [0, 0]) and np.allclose(xs[1], [1, 0])
        assert np.allclose(xs[2], x2, rtol=0, atol=1e-12)
        assert np.allclose(xs[3], x3, rtol=0, atol=1e-12)
        assert (st, it, ev) == ("converged", 3, 4)
