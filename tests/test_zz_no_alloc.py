"""sdot_eval_device_noalloc (the evaluation that allocates nothing; its
buffers come from the library's fixed 128 MiB device array) against
sdot_eval_device on the same inputs, through the C ABI.

Two calls in a row share the array, the second with a smaller input:
  1. 32^3 cells, QGauss(2) (262,144 q-points), 20,000 targets: the default
     segment counts need more than 128 MiB, so the call halves them;
     FP64 |F - F_ref| <= 1e-10 |F_ref|, max_j |g_j - g_ref_j| <= 1e-10 max_j
     (|g_ref_j| + nu_j); FP32_LSE |F - F_ref| <= 1e-8 |F_ref| and the gradient
     within 1e-6 of max(nu_j, m_j);
  2. 12^3 cells, 3,000 targets: the default segment counts fit, the plan is
     the one sdot_eval_device runs, and the results are bitwise equal.
Then an evaluation that does not fit with one segment per sweep is refused
before any launch, and the context's own evaluations are unaffected."""
import ctypes as C
import sys

import numpy as np
import pytest

import semidiscreteot_amd as sd
from semidiscreteot_amd import devmem
from semidiscreteot_amd import _lib as L
from tests.cases import structured_case

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-noalloc] {msg}", flush=True)
    sys.stdout.flush()


def make(src, tgt, precision, eps):
    s = sd.SotSolver(src.dim)
    s.setup_source(src.x, src.w, src.center, src.nq)
    s.setup_target(tgt.y, tgt.nu)
    s.configure(sd.SolverParameters(epsilon=eps, use_log_sum_exp_trick=True, precision=precision))
    return s


def run(s, R, psi, noalloc):
    d_psi = devmem.DeviceArray.from_host(psi)
    d_g = devmem.DeviceArray(psi.size)
    d_fc = devmem.DeviceArray(4)
    fn = s.lib.sdot_eval_device_noalloc if noalloc else s.lib.sdot_eval_device
    L.check(fn(s._h, float(R), C.c_void_p(d_psi.address), C.c_void_p(d_g.address),
               C.c_void_p(d_fc.address)))
    s.synchronize()
    fc = d_fc.download()
    return fc[0], d_g.download(), fc[1]


def compare(name, a, b, nu, precision, bitwise):
    (F1, g1, C1), (F0, g0, C0) = a, b
    relF = abs(F1 - F0) / abs(F0)
    if precision == "fp64":
        relg = np.max(np.abs(g1 - g0)) / np.max(np.abs(g0) + nu)
        say(f"{name}: F rel {relF:.3e}, grad rel {relg:.3e}, bitwise "
            f"{bool(F1 == F0 and np.array_equal(g1, g0))}")
        assert relF <= 1e-10 and relg <= 1e-10
        assert abs(C1 - C0) <= 1e-10 * abs(C0)
    else:
        relg = np.max(np.abs(g1 - g0) / np.maximum(nu, g0 + nu))
        say(f"{name}: F rel {relF:.3e}, grad / max(nu, m) {relg:.3e}")
        assert relF <= 1e-8 and relg <= 1e-6
    if bitwise:
        assert F1 == F0 and C1 == C0 and np.array_equal(g1, g0)


@pytest.mark.parametrize("precision", ["fp64", "fp32_lse"])
def test_noalloc_matches_eval_device_two_calls(precision):
    big = structured_case(3, 32, 2, 20000, seed=5, gmm_source=True, gmm_targets=True,
                          dirichlet=5)
    small = structured_case(3, 12, 2, 3000, seed=21, gmm_source=True, gmm_targets=True,
                            dirichlet=5)
    for name, (src, tgt), eps, R, bitwise in (
            ("32^3 x 20000 (segments halved)", big, 1e-3, 0.3, False),
            ("12^3 x 3000", small, 1e-3, 0.35, True)):
        psi = 0.002 * np.random.default_rng(3).standard_normal(tgt.n)
        say(f"{precision} {name}: noalloc")
        s = make(src, tgt, precision, eps)
        a = run(s, R, psi, True)
        s.close()
        say(f"{precision} {name}: eval_device on a fresh context")
        s = make(src, tgt, precision, eps)
        b = run(s, R, psi, False)
        s.close()
        compare(name, a, b, tgt.nu, precision, bitwise)


def test_noalloc_refuses_oversize_and_leaves_context_state():
    src, tgt = structured_case(3, 8, 2, 3000, seed=2, gmm_source=True)
    s = make(src, tgt, "fp32_lse", 1e-2)
    psi = 0.01 * np.random.default_rng(4).standard_normal(tgt.n)
    psi2 = psi + 1e-4 * np.random.default_rng(5).standard_normal(tgt.n)
    # the context's seeded FP32 history survives a no-allocation call between
    # two of its own evaluations
    ref1 = run(s, 0.4, psi, False)
    na = run(s, 0.4, psi2, True)
    ref2 = run(s, 0.4, psi2, False)
    s.close()
    s = make(src, tgt, "fp32_lse", 1e-2)
    run(s, 0.4, psi, False)
    ref2_clean = run(s, 0.4, psi2, False)
    s.close()
    say("history kept across a no-allocation call")
    assert ref2[0] == ref2_clean[0] and np.array_equal(ref2[1], ref2_clean[1])
    compare("noalloc vs seeded", na, ref2, tgt.nu, "fp32_lse", False)
    del ref1
    # 900,000 targets need about 144 MB with one segment per sweep
    n = 900_000
    rng = np.random.default_rng(6)
    y = rng.random((n, 3)).ravel()
    nu = np.full(n, 1.0 / n)
    s = sd.SotSolver(3)
    s.setup_source(src.x, src.w, src.center, src.nq)
    s.setup_target(y, nu)
    s.configure(sd.SolverParameters(epsilon=1e-2, use_log_sum_exp_trick=True, precision="fp64"))
    d_psi = devmem.DeviceArray.from_host(np.zeros(n))
    d_g = devmem.DeviceArray(n)
    d_fc = devmem.DeviceArray(4)
    rc = s.lib.sdot_eval_device_noalloc(s._h, 0.2, C.c_void_p(d_psi.address),
                                         C.c_void_p(d_g.address), C.c_void_p(d_fc.address))
    msg = s.lib.sdot_last_error().decode()
    say(f"oversize: rc {rc}: {msg}")
    assert rc == -1 and "no-allocation entry point holds 134217728" in msg
    s.close()
