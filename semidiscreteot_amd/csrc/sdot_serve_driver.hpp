// sdot_serve_driver.hpp -- host side of one served solve: the loop of
// lbfgs_solve (sdot_bfgs.hpp; SotSolver::solve, SotSolver.cc:105-215) cut
// into steps.  Each step the device performs the action next_in() asks for
// and one evaluation at the resulting point, then on_result() consumes what
// the device measured:
//   kActNone   evaluate at the current point (first evaluation, or a repeat),
//   kActIter   store (s, y) if s.y > 0, two-loop direction p = -H g, first
//              trial step a1, x = x0 + a1 p (x0 = the accepted point),
//   kActTrial  x = x0 + alpha p for the line search's next trial.
// Precision escalation (optional): a solve that starts with low-precision
// evaluations asks to be switched to FP64 (on_result returns kSwitch) once its
// stopping value is within `polish` times the tolerance, or when a line search
// fails; the caller switches the evaluation mode, and the next step evaluates
// again at the current iterate (x0 after a failed search, restored by a trial
// with alpha = 0) and restarts the iteration from there (first trial step 1,
// the stored curvature pairs kept).  The stopping decision is thus always
// taken on an FP64 evaluation.
// Plain C++ (no HIP): shared by the serving queue and its CPU check.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "sdot_bfgs.hpp"

namespace sdot {

constexpr int kHist = 5;  // L-BFGS history (deal.II SolverBFGS default)
enum : int { kActNone = 0, kActIter = 1, kActTrial = 2 };

// Written by the host before a step.  The noise-aware line search is on for
// the new direction when surrogate != 0 or |g0| < noise (first trial step 1);
// noise = 0 and surrogate = 0 select the reference search.  h0 > 0: initial
// inverse Hessian diag(h0 / nu_j) of the two-loop recursion (0: identity).
struct SlotIn {
  int32_t action, first, has_prev, surrogate;
  double alpha, f, f_prev, tol, noise, h0;
};

// Written by the device during a step (F and C first: the final reduction
// writes them as a pair).  g0 / a1: slope along the new direction and first
// trial step (kActIter).  lost > 0: a history-seeded FP32 sweep lost range
// and the evaluation must be repeated unseeded.
struct SlotOut {
  double F, C, gp, l1, cw, relerr, g0, a1, lost;
};

class SolveDriver {
 public:
  enum Phase { kStart = 0, kNewDir = 1, kTrial = 2, kRestart = 3 };
  static constexpr int kSwitch = -2;

  // noise_rel > 0: noise-aware line search with noise floor noise_rel |f|.
  // polish > 0: low-precision start, escalation to FP64 (noise floor
  // noise_rel_fp64 from then on) as described above.
  void init(int control, double tol, uint32_t max_it, double noise_rel = 0.0,
            double polish = 0.0, double noise_rel_fp64 = 0.0) {
    control_ = control;
    tol_ = tol;
    max_it_ = max_it;
    noise_rel_ = noise_rel;
    polish_ = polish;
    low_ = polish > 0.0;
    noise_hi_ = noise_rel > 0.0 ? noise_rel_fp64 : 0.0;
  }

  // True while the solve evaluates in low precision.
  bool low_precision() const { return low_; }

  SlotIn next_in() const {
    SlotIn c{};
    c.action = kActNone;
    if (!redo_) {
      if (phase_ == kNewDir && !dir_ready_) c.action = kActIter;
      if (phase_ == kTrial) c.action = kActTrial;
      if (phase_ == kRestart && restore_) c.action = kActTrial;  // alpha = 0: x = x0
    }
    c.first = first_ ? 1 : 0;
    c.has_prev = first_ ? 0 : 1;
    c.alpha = phase_ == kTrial ? ls_.trial() : 0.0;
    c.f = f_;
    c.f_prev = f_prev_;
    c.tol = tol_;
    c.noise = noise_rel_ > 0.0 ? 100.0 * noise_rel_ * std::fabs(f_) : 0.0;
    c.surrogate = sur_ ? 1 : 0;
    return c;
  }

  // Returns a sdot_solve_status (>= 0) when the solve has finished, kSwitch
  // when it must continue in FP64, -1 when it continues.  repeat: the
  // evaluation lost range and is done again.
  int on_result(const SlotOut &o, bool repeat) {
    if (repeat) {
      if (phase_ == kNewDir && !dir_ready_) {
        g0_ = o.g0;
        a1_ = o.a1;
        dir_ready_ = true;
      }
      redo_ = true;
      return -1;
    }
    redo_ = false;
    ++evals_;
    relerr_ = o.relerr;
    int r;
    if (phase_ == kStart || phase_ == kRestart) {
      f_ = o.F;
      return after_point(o);
    }
    if (phase_ == kNewDir) {
      const double g0 = dir_ready_ ? g0_ : o.g0;
      const double a1 = dir_ready_ ? a1_ : o.a1;
      dir_ready_ = false;
      // the device chose a1 with the same test (first_step)
      const double noise = noise_rel_ > 0.0 ? 100.0 * noise_rel_ * std::fabs(f_) : 0.0;
      sur_ = noise_rel_ > 0.0 && (sur_ || std::fabs(g0) < noise);
      ls_.start(f_, g0, a1, sur_);
      f_prev_ = f_;
    }
    r = ls_.feed(o.F, o.gp);
    if (r == LineSearchSM::kEval) {
      phase_ = kTrial;
      return -1;
    }
    if (r == LineSearchSM::kFail) {
      ++k_;
      if (low_) return escalate(true);
      f_ = o.F;
      return 2;  // SDOT_SOLVE_LINESEARCH_FAIL
    }
    f_ = o.F;
    // accepted: the point just evaluated is the new iterate
    first_ = false;
    ++k_;
    return after_point(o);
  }

  int iterations() const { return k_; }
  int evaluations() const { return evals_; }
  double functional() const { return f_; }
  double check_value() const { return check_; }
  double relerr() const { return relerr_; }

 private:
  // Stopping test of SotSolver::solve (SotSolver.cc:156-162,
  // SotSolver.h:382-446) at the current iterate.
  int after_point(const SlotOut &o) {
    bool ok;
    if (control_ == 1) {  // componentwise: max_j (|g_j| - nu_j tol) <= DBL_MIN
      check_ = o.cw;
      ok = check_ <= DBL_MIN;
    } else {              // l1norm: sum_j |g_j| <= tol
      check_ = o.l1;
      ok = check_ <= tol_;
    }
    if (low_ && !std::isnan(check_)) {
      // within reach of the tolerance: finish in FP64
      const double v = control_ == 1 ? o.relerr : o.l1;
      if (v <= polish_ * tol_) return escalate(false);
    }
    if (ok) return 0;                                   // SDOT_SOLVE_CONVERGED
    if (std::isnan(check_)) return 3;                   // SDOT_SOLVE_NONFINITE
    if ((uint32_t)k_ >= max_it_) return 1;              // SDOT_SOLVE_MAX_ITER
    phase_ = kNewDir;
    return -1;
  }

  int escalate(bool restore) {
    low_ = false;
    restore_ = restore;
    phase_ = kRestart;
    first_ = true;
    sur_ = false;
    noise_rel_ = noise_hi_;
    return kSwitch;
  }

  int control_ = 0;
  double tol_ = 0.0, noise_rel_ = 0.0, polish_ = 0.0, noise_hi_ = 0.0;
  bool low_ = false, restore_ = false;
  uint32_t max_it_ = 0;
  Phase phase_ = kStart;
  LineSearchSM ls_;
  double f_ = 0.0, f_prev_ = 0.0, check_ = 0.0, relerr_ = 0.0;
  bool first_ = true, redo_ = false, dir_ready_ = false, sur_ = false;
  double g0_ = 0.0, a1_ = 0.0;
  int k_ = 0, evals_ = 0;
};

}  // namespace sdot
