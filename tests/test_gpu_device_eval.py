"""The device-pointer entry point sdot_eval_device (psi, grad and (F, C) in
HBM, asynchronous) against the FP64 oracle, in the FP32_LSE mode whose sweep
seeds each work item from the previous evaluation's maxima.

Tolerances: relative marginal error max_j |g_j - g_ref_j| / max(nu_j, m_j)
<= 1e-6 (m_j = g_ref_j + nu_j, the assembled marginal) and
|F - F_ref| <= 1e-8 |F_ref|."""
import os
import sys

import numpy as np
import pytest

import semidiscreteot_amd as sd
from semidiscreteot_amd import devmem
from oracle import oracle as O
from tests.cases import structured_case

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu] {msg}", flush=True)
    sys.stdout.flush()


def case():
    src, tgt = structured_case(3, 12, 2, 3000, seed=21, gmm_source=True, gmm_targets=True,
                               dirichlet=5)
    return src, tgt


def make(src, tgt):
    s = sd.SotSolver(3)
    s.setup_source(src.x, src.w, src.center, src.nq)
    s.setup_target(tgt.y, tgt.nu)
    s.configure(sd.SolverParameters(epsilon=1e-3, use_log_sum_exp_trick=True,
                                    precision="fp32_lse"))
    return s


def device_eval(s, R, psi):
    d_psi = devmem.DeviceArray.from_host(psi)
    d_g = devmem.DeviceArray(psi.size)
    d_fc = devmem.DeviceArray(4)
    s.eval_device(R, d_psi.address, d_g.address, d_fc.address)
    s.synchronize()
    g = d_g.download()
    fc = d_fc.download()
    return fc[0], g, fc[1]


def check(src, tgt, R, psi, F, g):
    Fo, go, _ = O.evaluate_R(src, tgt, O.params(epsilon=1e-3, use_lse=1), R, psi)
    relm = np.max(np.abs(g - go) / np.maximum(tgt.nu, go + tgt.nu))
    relF = abs(F - Fo) / abs(Fo)
    say(f"error / max(nu, m) {relm:.3e}, F rel {relF:.3e}")
    assert relm <= 1e-6
    assert relF <= 1e-8


def test_device_entry_point_across_a_large_potential_jump():
    src, tgt = case()
    R = 0.35
    rng = np.random.default_rng(5)
    psi_a = 0.02 * rng.standard_normal(tgt.n)
    s = make(src, tgt)
    say("device eval, cold")
    F, g, _ = device_eval(s, R, psi_a)
    check(src, tgt, R, psi_a, F, g)
    # seeded next evaluation (small change), then a jump of ~30 log2 units in
    # the exponents: the history seed must not be used across it
    for step in (2e-4, 2e-2):
        psi_b = psi_a + step * rng.standard_normal(tgt.n)
        say(f"device eval after a potential change of {step:g}")
        F, g, _ = device_eval(s, R, psi_b)
        check(src, tgt, R, psi_b, F, g)
        psi_a = psi_b
    s.close()


def test_item_with_a_seed_too_high_is_swept_again():
    """A seeded reference value far above an item's new maximum (forced through
    the SDOT_TEST_SEED_BIAS hook: the seed bound itself never lies that high)
    is detected in the sweep, which sweeps the item again against its exact
    maximum: the result equals a cold evaluation bitwise."""
    src, tgt = case()
    R = 0.35
    rng = np.random.default_rng(6)
    psi_a = 0.02 * rng.standard_normal(tgt.n)
    psi_b = psi_a + 1e-4 * rng.standard_normal(tgt.n)
    os.environ["SDOT_TEST_SEED_BIAS"] = "40"
    try:
        s = make(src, tgt)
    finally:
        del os.environ["SDOT_TEST_SEED_BIAS"]
    say("biased seed: first evaluation (cold), second (seeded 40 log2 units too high)")
    device_eval(s, R, psi_a)
    F1, g1, C1 = device_eval(s, R, psi_b)
    cold = make(src, tgt)
    F2, g2, C2 = device_eval(cold, R, psi_b)
    say(f"re-swept vs cold: F {F1!r} {F2!r}")
    assert F1 == F2 and C1 == C2 and np.array_equal(g1, g2)
    check(src, tgt, R, psi_b, F1, g1)
    # the host entry point takes the same path
    F3, g3, _ = s.evaluate_functional_R(R, psi_b + 1e-5)
    F4, g4, _ = cold.evaluate_functional_R(R, psi_b + 1e-5)
    assert abs(F3 - F4) <= 1e-8 * abs(F4)
    s.close()
    cold.close()
