// sdot_bfgs.hpp -- FP64 limited-memory BFGS with the line search that
// SotSolver::solve uses through deal.II (SotSolver.cc:190-198):
//   * SolverBFGS: history of 5 (s, y) pairs, newest first, no preconditioner,
//     pair stored only when s.y > 0; the in-tree text of the algorithm is
//     core/wasserstein_solver_bfgs.h:402-491.
//   * default line minimisation: first trial step 1 on the first iteration,
//     else min(1, -2.02 * max(f_prev - f, 100 eps_mach) / g0)
//     (wasserstein_solver_bfgs.h:322-365); Wolfe parameters eta = 0.9,
//     mu = 0.001, at most 20 evaluations.
//   * LineMinimization::line_search (line_minimization.h:498-749) with
//     poly_fit (cubic, then quadratic, then bisection of the safeguarded
//     interval, line_minimization.h:423-455): bracketing growth 9, sectioning
//     safeguards 0.1 / 0.5.  Running out of evaluations is reported as
//     kLineSearchFail (stock deal.II throws there).
//   * optional noise-aware mode (not in the reference): once the decrease the
//     first trial predicts, |g0|, falls below 100 times the noise floor of the
//     functional (noise_rel |f0|), every later line search of the solve
//     compares functional differences taken from the trapezoid integral of the
//     measured slopes, f(a) ~ f(b) + (a - b) (g(a) + g(b)) / 2 from the nearest
//     evaluated point b, and starts from the step 1 (the measured differences
//     of f are noise from then on, also for the first-step rule).  This is the idea of the approximate
//     Wolfe conditions of Hager & Zhang (SIAM J. Optim. 16 (2005) 170-192):
//     near the minimiser F changes by less than its rounding noise while the
//     gradient is still accurate, and the integral is exact on quadratics.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <type_traits>
#include <vector>

namespace sdot {

enum BfgsStatus { kConverged = 0, kMaxIter = 1, kLineSearchFail = 2, kNonFinite = 3 };
enum CheckState { kIterate = 0, kSuccess = 1, kFailure = 2 };

struct BfgsResult {
  int status = kMaxIter;
  int iterations = 0;
  int evaluations = 0;
  double f = 0.0;
  double check = 0.0;
};

inline double bfgs_dot(const double *a, const double *b, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// One-dimensional Wolfe search along p from x0.  phi(a, f, g) evaluates the
// objective at x0 + a p and returns f and the slope g = grad . p.
template <class Phi>
bool wolfe_line_search(Phi &&phi, double f0, double g0, double a1, double &alpha,
                       bool surrogate = false) {
  constexpr double eta = 0.9, mu = 0.001, a_max = DBL_MAX;
  constexpr unsigned max_evals = 20;
  constexpr double grow = 9.0, shrink_lo = 0.1, shrink_hi = 0.5;
  // in the noise-aware regime every value is an offset from f0 (its
  // increments would be absorbed by the magnitude of f0 otherwise)
  const double fb = surrogate ? 0.0 : f0;
  const double slope_tol = eta * std::fabs(g0);
  const double f_floor = fb + a_max * mu * g0;
  auto armijo = [&](double a, double f) { return f <= fb + a * mu * g0; };
  auto curvature = [&](double g) { return std::fabs(g) <= slope_tol; };

  double lo_a = NAN, lo_f = NAN, lo_g = NAN, hi_a = NAN, hi_f = NAN, hi_g = NAN;
  unsigned used = 0;
  double prev_a = 0.0, prev_f = fb, prev_g = g0;
  double a = a1, f = 0.0, g = 0.0;
  bool bracketed = false;
  while (used < max_evals) {
    phi(a, f, g);
    ++used;
    if (surrogate) f = prev_f + (a - prev_a) * (0.5 * (prev_g + g));
    if (f <= f_floor || a == a_max) { alpha = a; return true; }
    if (!armijo(a, f) || (f >= prev_f && used > 1)) {
      lo_a = prev_a; lo_f = prev_f; lo_g = prev_g;
      hi_a = a; hi_f = f; hi_g = g;
      bracketed = true;
      break;
    }
    if (curvature(g)) { alpha = a; return true; }
    if (g >= 0) {
      lo_a = a; lo_f = f; lo_g = g;
      hi_a = prev_a; hi_f = prev_f; hi_g = prev_g;
      bracketed = true;
      break;
    }
    const double next = std::fmin(a_max, a + grow * (a - prev_a));
    prev_a = a; prev_f = f; prev_g = g;
    a = next;
  }
  if (!bracketed || used >= max_evals) return false;

  while (used < max_evals) {
    const double s1 = lo_a + shrink_lo * (hi_a - lo_a);
    const double s2 = hi_a - shrink_hi * (hi_a - lo_a);
    const double bl = std::fmin(s1, s2), bh = std::fmax(s1, s2);
    // cubic through both ends with slopes
    double trial = 0.0;
    bool ok = false;
    {
      const double b1 = lo_g + hi_g - 3.0 * (lo_f - hi_f) / (lo_a - hi_a);
      const double disc = b1 * b1 - lo_g * hi_g;
      if (disc >= 0) {
        const double b2 = std::sqrt(disc);
        const double den = lo_a < hi_a ? hi_g - lo_g + 2.0 * b2 : lo_g - hi_g + 2.0 * b2;
        if (den != 0.0) {
          trial = lo_a < hi_a ? hi_a - (hi_a - lo_a) * (hi_g + b2 - b1) / den
                              : lo_a - (lo_a - hi_a) * (lo_g + b2 - b1) / den;
          ok = trial >= bl && trial <= bh;
        }
      }
    }
    if (!ok) {  // quadratic through lo (value, slope) and hi (value)
      const double den = 2.0 * lo_g * hi_a - 2.0 * lo_g * lo_a - 2.0 * hi_f + 2.0 * lo_f;
      if (den != 0) {
        trial = (lo_g * (hi_a * hi_a - lo_a * lo_a) + 2.0 * (lo_f - hi_f) * lo_a) / den;
        ok = trial >= bl && trial <= bh;
      }
    }
    if (!ok) trial = 0.5 * (bl + bh);
    a = trial;
    phi(a, f, g);
    ++used;
    if (surrogate) f = lo_f + (a - lo_a) * (0.5 * (lo_g + g));
    if (!armijo(a, f) || f >= lo_f) {
      hi_a = a; hi_f = f; hi_g = g;
    } else {
      if (curvature(g)) { alpha = a; return true; }
      if (g * (hi_a - lo_a) >= 0) {
        hi_a = lo_a; hi_f = lo_f; hi_g = lo_g;
      }
      lo_a = a; lo_f = f; lo_g = g;
    }
  }
  return false;
}

// eval(x, g) -> f ; check(step, g, &value) -> CheckState
// First trial step (wasserstein_solver_bfgs.h:335-341), or 1 in the
// noise-aware regime.
inline double first_step(bool first, bool surrogate, double f_prev, double f, double g0) {
  if (first || surrogate) return 1.0;
  const double df = std::fmax(f_prev - f, 100.0 * DBL_EPSILON);
  return std::fmin(1.0, -1.01 * 2.0 * df / g0);
}

// The vectors of lbfgs_core live in a Space, which holds the work vectors
// vec(kG .. kX0) and the (s, y) pairs and does the vector arithmetic:
//   copy(d, s)               d = s
//   dot(a, b)                a . b
//   lin(x, x0, a, p)         x = x0 + a p
//   diff(s, x, y, g)         s = x - s, y = g - y
//   direction(g, p, slots, rho, count)
//                            p = -H g, H over the pairs slots[0..count) (newest
//                            first) with their rho, on the initial inverse
//                            Hessian of the space (identity or a diagonal)
//   store(slot)              pair `slot` takes vec(kS), vec(kY) (whose contents
//                            are free afterwards)
//   ok()                     false after a failure of the space's own (which
//                            ends the solve with kNonFinite)
// x is a vector of the space; eval(x, g) and check(k, g, cv) take the space's
// vectors.
enum { kVecG = 0, kVecP, kVecS, kVecY, kVecX0, kNumVecs };

template <class Space, class Eval, class Check>
BfgsResult lbfgs_core(Space &sp, double *x, Eval &&eval, Check &&check, int history,
                      double noise_rel) {
  BfgsResult res;
  std::vector<double> rho(history), coef(history);
  std::vector<int> slots(history);
  int count = 0, newest = -1;  // ring buffer; slot of pair i (0 = newest) is (newest - i) mod h
  auto slot = [&](int i) { return ((newest - i) % history + history) % history; };
  auto stop = [&](int status, int k, double f, double cv) {
    res.status = status;
    res.iterations = k;
    res.f = f;
    res.check = cv;
    return res;
  };
  double *g = sp.vec(kVecG), *p = sp.vec(kVecP), *x0 = sp.vec(kVecX0);

  double f = eval(x, g);
  res.evaluations = 1;
  int k = 0;
  double cv = 0.0;
  int st = check(k, g, cv);
  bool first = true, sur = false;
  double f_prev = 0.0;
  while (st == kIterate) {
    // two-loop recursion, H0 = identity or diag(h0)
    for (int i = 0; i < count; ++i) {
      slots[i] = slot(i);
      coef[i] = rho[slot(i)];
    }
    sp.direction(g, p, slots.data(), coef.data(), count);
    double *sk = sp.vec(kVecS), *yk = sp.vec(kVecY);
    sp.copy(sk, x);
    sp.copy(yk, g);

    const double f0 = f, g0 = sp.dot(g, p);
    if (!sp.ok()) return stop(kNonFinite, k, f, cv);
    sp.copy(x0, x);
    sur = noise_rel > 0.0 && (sur || std::fabs(g0) < 100.0 * noise_rel * std::fabs(f0));
    const double a1 = first_step(first, sur, f_prev, f, g0);
    f_prev = f;
    auto phi = [&](double a, double &fo, double &go) {
      sp.lin(x, x0, a, p);
      f = eval(x, g);
      ++res.evaluations;
      fo = f;
      go = sp.dot(g, p);
    };
    double alpha = 0.0;
    if (!wolfe_line_search(phi, f0, g0, a1, alpha, sur))
      return stop(sp.ok() ? kLineSearchFail : kNonFinite, k + 1, f, cv);
    first = false;
    sp.diff(sk, x, yk, g);
    ++k;
    st = check(k, g, cv);
    if (st != kIterate) break;
    const double curv = sp.dot(sk, yk);
    if (!sp.ok()) return stop(kNonFinite, k, f, cv);
    if (curv > 0.0 && history > 0) {
      newest = (newest + 1) % history;
      sp.store(newest);
      rho[newest] = 1.0 / curv;
      if (count < history) ++count;
    }
  }
  res.status = st == kSuccess ? kConverged : (std::isnan(cv) ? kNonFinite : kMaxIter);
  res.iterations = k;
  res.f = f;
  res.check = cv;
  return res;
}

// Host vectors of length n; the initial inverse Hessian is diag(h0), or the
// identity when h0 == nullptr (deal.II's SolverBFGS).
class HostSpace {
 public:
  HostSpace(int64_t n, int history, const double *h0)
      : n_(n), h0_(h0), work_(kNumVecs, std::vector<double>(n)),
        S_(history, std::vector<double>(n)), Y_(history, std::vector<double>(n)),
        coef_(history) {}
  double *vec(int i) { return work_[i].data(); }
  bool ok() const { return true; }
  void copy(double *d, const double *s) { std::memcpy(d, s, sizeof(double) * n_); }
  double dot(const double *a, const double *b) { return bfgs_dot(a, b, n_); }
  void lin(double *x, const double *x0, double a, const double *p) {
    for (int64_t t = 0; t < n_; ++t) x[t] = x0[t] + a * p[t];
  }
  void diff(double *s, const double *x, double *y, const double *g) {
    for (int64_t t = 0; t < n_; ++t) {
      s[t] = x[t] - s[t];
      y[t] = g[t] - y[t];
    }
  }
  void store(int slot) {
    S_[slot].swap(work_[kVecS]);
    Y_[slot].swap(work_[kVecY]);
  }
  void direction(const double *g, double *p, const int *slots, const double *rho, int count) {
    const int64_t n = n_;
    std::memcpy(p, g, sizeof(double) * n);
    for (int i = 0; i < count; ++i) {
      const int s = slots[i];
      coef_[i] = rho[i] * bfgs_dot(S_[s].data(), p, n);
      const double c = -coef_[i];
      const double *ys = Y_[s].data();
      for (int64_t t = 0; t < n; ++t) p[t] += c * ys[t];
    }
    if (h0_)
      for (int64_t t = 0; t < n; ++t) p[t] *= h0_[t];
    for (int i = count - 1; i >= 0; --i) {
      const int s = slots[i];
      const double c2 = rho[i] * bfgs_dot(Y_[s].data(), p, n);
      const double c = coef_[i] - c2;
      const double *ss = S_[s].data();
      for (int64_t t = 0; t < n; ++t) p[t] += c * ss[t];
    }
    for (int64_t t = 0; t < n; ++t) p[t] = -p[t];
  }

 private:
  int64_t n_;
  const double *h0_;
  std::vector<std::vector<double>> work_, S_, Y_;
  std::vector<double> coef_;
};

// noise_rel > 0 selects the noise-aware line search with noise floor
// noise_rel |f|; 0 is the reference algorithm.
// h0: diagonal of the initial inverse Hessian of the two-loop recursion
// (nullptr: the identity, as deal.II's SolverBFGS).
template <class Eval, class Check>
BfgsResult lbfgs_solve(int64_t n, double *x, Eval &&eval, Check &&check, int history = 5,
                       double noise_rel = 0.0, const double *h0 = nullptr) {
  HostSpace sp(n, history, h0);
  return lbfgs_core(sp, x, eval, check, history, noise_rel);
}

}  // namespace sdot

namespace sdot {

// Resumable form of wolfe_line_search for drivers that evaluate many line
// searches in lock-step (one batched evaluation per step): the caller
// evaluates at trial() and feeds (f, slope) back.  The decisions are those of
// wolfe_line_search, step for step.
class LineSearchSM {
 public:
  enum { kEval = 0, kAccept = 1, kFail = -1 };

  void start(double f0, double g0, double a1, bool surrogate = false) {
    sur_ = surrogate;
    f0_ = surrogate ? 0.0 : f0;  // noise-aware: offsets from f0
    g0_ = g0;
    tol_ = 0.9 * std::fabs(g0);
    floor_ = f0_ + DBL_MAX * 0.001 * g0;
    used_ = 0;
    phase_ = 0;
    prev_a_ = 0.0;
    prev_f_ = f0_;
    prev_g_ = g0;
    a_ = a1;
  }
  double trial() const { return a_; }
  unsigned evaluations() const { return used_; }

  int feed(double f, double g) {
    constexpr unsigned max_evals = 20;
    ++used_;
    const double a = a_;
    if (sur_)
      f = phase_ == 0 ? prev_f_ + (a - prev_a_) * (0.5 * (prev_g_ + g))
                      : lo_f_ + (a - lo_a_) * (0.5 * (lo_g_ + g));
    if (phase_ == 0) {
      if (f <= floor_ || a == DBL_MAX) return kAccept;
      if (!armijo(a, f) || (f >= prev_f_ && used_ > 1)) {
        lo_a_ = prev_a_; lo_f_ = prev_f_; lo_g_ = prev_g_;
        hi_a_ = a; hi_f_ = f; hi_g_ = g;
        return to_sectioning(max_evals);
      }
      if (std::fabs(g) <= tol_) return kAccept;
      if (g >= 0) {
        lo_a_ = a; lo_f_ = f; lo_g_ = g;
        hi_a_ = prev_a_; hi_f_ = prev_f_; hi_g_ = prev_g_;
        return to_sectioning(max_evals);
      }
      const double next = std::fmin(DBL_MAX, a + 9.0 * (a - prev_a_));
      prev_a_ = a; prev_f_ = f; prev_g_ = g;
      a_ = next;
      return used_ < max_evals ? kEval : kFail;
    }
    if (!armijo(a, f) || f >= lo_f_) {
      hi_a_ = a; hi_f_ = f; hi_g_ = g;
    } else {
      if (std::fabs(g) <= tol_) return kAccept;
      if (g * (hi_a_ - lo_a_) >= 0) {
        hi_a_ = lo_a_; hi_f_ = lo_f_; hi_g_ = lo_g_;
      }
      lo_a_ = a; lo_f_ = f; lo_g_ = g;
    }
    if (used_ >= max_evals) return kFail;
    a_ = section_trial();
    return kEval;
  }

 private:
  bool armijo(double a, double f) const { return f <= f0_ + a * 0.001 * g0_; }
  int to_sectioning(unsigned max_evals) {
    phase_ = 1;
    if (used_ >= max_evals) return kFail;
    a_ = section_trial();
    return kEval;
  }
  double section_trial() const {
    const double s1 = lo_a_ + 0.1 * (hi_a_ - lo_a_);
    const double s2 = hi_a_ - 0.5 * (hi_a_ - lo_a_);
    const double bl = std::fmin(s1, s2), bh = std::fmax(s1, s2);
    const double b1 = lo_g_ + hi_g_ - 3.0 * (lo_f_ - hi_f_) / (lo_a_ - hi_a_);
    const double disc = b1 * b1 - lo_g_ * hi_g_;
    if (disc >= 0) {
      const double b2 = std::sqrt(disc);
      const double den = lo_a_ < hi_a_ ? hi_g_ - lo_g_ + 2.0 * b2 : lo_g_ - hi_g_ + 2.0 * b2;
      if (den != 0.0) {
        const double t = lo_a_ < hi_a_ ? hi_a_ - (hi_a_ - lo_a_) * (hi_g_ + b2 - b1) / den
                                       : lo_a_ - (lo_a_ - hi_a_) * (lo_g_ + b2 - b1) / den;
        if (t >= bl && t <= bh) return t;
      }
    }
    const double den = 2.0 * lo_g_ * hi_a_ - 2.0 * lo_g_ * lo_a_ - 2.0 * hi_f_ + 2.0 * lo_f_;
    if (den != 0) {
      const double t = (lo_g_ * (hi_a_ * hi_a_ - lo_a_ * lo_a_) + 2.0 * (lo_f_ - hi_f_) * lo_a_) / den;
      if (t >= bl && t <= bh) return t;
    }
    return 0.5 * (bl + bh);
  }
  bool sur_ = false;
  double f0_ = 0, g0_ = 0, tol_ = 0, floor_ = 0;
  unsigned used_ = 0;
  int phase_ = 0;
  double a_ = 0, prev_a_ = 0, prev_f_ = 0, prev_g_ = 0;
  double lo_a_ = 0, lo_f_ = 0, lo_g_ = 0, hi_a_ = 0, hi_f_ = 0, hi_g_ = 0;
};

}  // namespace sdot
