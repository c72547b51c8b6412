// Drives the served-solve state machine (SolveDriver, sdot_serve_driver.hpp)
// with a host emulation of the device actions (history update, two-loop
// direction, trial points, slope and stopping statistics, in the order the
// loop form computes them) and compares with lbfgs_solve on Rosenbrock:
// iterates, iteration and evaluation counts must be identical.  A second run
// flags every 5th evaluation as "lost" (repeated unseeded) and must give the
// same result.  With a diagonal initial inverse Hessian the two forms agree
// the same way, and on a separable quadratic whose Hessian it inverts the
// first step reaches the minimiser.
#include <cmath>
#include <cstdio>
#include <vector>

#include "../semidiscreteot_amd/csrc/sdot_serve_driver.hpp"
using namespace sdot;

static double rosen(const double *x, double *g, int64_t n) {
  double f = 0;
  for (int64_t i = 0; i < n; ++i) g[i] = 0;
  for (int64_t i = 0; i + 1 < n; ++i) {
    const double a = 1 - x[i], b = x[i + 1] - x[i] * x[i];
    f += a * a + 100 * b * b;
    g[i] += -2 * a - 400 * x[i] * b;
    g[i + 1] += 200 * b;
  }
  return f;
}

struct Emu {
  int64_t n;
  std::vector<double> x, g, x0, gv, p, nu;
  std::vector<std::vector<double>> S, Y;
  double rho[kHist] = {0, 0, 0, 0, 0};
  int count = 0, newest = -1;
  explicit Emu(int64_t n_) : n(n_), x(n_), g(n_), x0(n_), gv(n_), p(n_), nu(n_, 1.0),
                             S(kHist, std::vector<double>(n_)), Y(kHist, std::vector<double>(n_)) {}
  double dot(const std::vector<double> &a, const std::vector<double> &b) const {
    double s = 0;
    for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
  }
  void action(const SlotIn &c, SlotOut &o) {
    if (c.action == kActTrial) {
      for (int64_t i = 0; i < n; ++i) x[i] = x0[i] + c.alpha * p[i];
      return;
    }
    if (c.action != kActIter) return;
    if (c.has_prev) {
      double curv = 0;
      for (int64_t i = 0; i < n; ++i) curv += (x[i] - x0[i]) * (g[i] - gv[i]);
      if (curv > 0) {
        const int sl = (newest + 1) % kHist;
        for (int64_t i = 0; i < n; ++i) {
          S[sl][i] = x[i] - x0[i];
          Y[sl][i] = g[i] - gv[i];
        }
        rho[sl] = 1.0 / curv;
        newest = sl;
        count = count < kHist ? count + 1 : kHist;
      }
    }
    p = g;
    double c1[kHist];
    for (int i = 0; i < count; ++i) {
      const int sl = ((newest - i) % kHist + kHist) % kHist;
      c1[i] = rho[sl] * dot(S[sl], p);
      for (int64_t t = 0; t < n; ++t) p[t] += -c1[i] * Y[sl][t];
    }
    if (c.h0 > 0.0)
      for (int64_t t = 0; t < n; ++t) p[t] *= c.h0 / nu[t];
    for (int i = count - 1; i >= 0; --i) {
      const int sl = ((newest - i) % kHist + kHist) % kHist;
      const double cc = c1[i] - rho[sl] * dot(Y[sl], p);
      for (int64_t t = 0; t < n; ++t) p[t] += cc * S[sl][t];
    }
    for (auto &v : p) v = -v;
    const double g0 = dot(g, p);
    const double df = std::fmax(c.f_prev - c.f, 100.0 * 2.220446049250313e-16);
    const double a1 = (c.first || c.surrogate || std::fabs(g0) < c.noise)
                          ? 1.0 : std::fmin(1.0, -1.01 * 2.0 * df / g0);
    x0 = x;
    gv = g;
    for (int64_t i = 0; i < n; ++i) x[i] = x0[i] + a1 * p[i];
    o.g0 = g0;
    o.a1 = a1;
  }
  void evaluate(SlotOut &o, double tol) {
    o.F = rosen(x.data(), g.data(), n);
    o.gp = dot(g, p);
    o.l1 = 0;
    o.cw = -INFINITY;
    o.relerr = 0;
    for (int64_t i = 0; i < n; ++i) {
      o.l1 += std::fabs(g[i]);
      o.cw = std::fmax(o.cw, std::fabs(g[i]) - nu[i] * tol);
      o.relerr = std::fmax(o.relerr, std::fabs(g[i]) / nu[i]);
    }
  }
};

static int served(int64_t n, double tol, uint32_t max_it, int lose_every, std::vector<double> &xout,
                  int &it, int &ev, double noise_rel = 0.0, double h0 = 0.0,
                  const std::vector<double> *nu = nullptr) {
  Emu e(n);
  for (auto &v : e.x) v = -1.0;
  e.x[1] = 1.2;
  if (nu) e.nu = *nu;
  SolveDriver d;
  d.init(0, tol, max_it, noise_rel);
  int status = -1, raw = 0;
  while (status < 0) {
    SlotIn c = d.next_in();
    c.h0 = h0;
    SlotOut o{};
    e.action(c, o);
    e.evaluate(o, tol);
    ++raw;
    const bool lost = lose_every > 0 && raw % lose_every == 0;
    status = d.on_result(o, lost);
  }
  xout = e.x;
  it = d.iterations();
  ev = d.evaluations();
  return status;
}

int main() {
  const int64_t n = 6;
  const double tol = 1e-9;
  const uint32_t max_it = 500;
  std::vector<double> xa(n, -1.0);
  xa[1] = 1.2;
  auto check = [&](int step, const double *g, double &cv) {
    cv = 0;
    for (int64_t i = 0; i < n; ++i) cv += std::fabs(g[i]);
    if (cv <= tol) return (int)kSuccess;
    if ((uint32_t)step >= max_it || std::isnan(cv)) return (int)kFailure;
    return (int)kIterate;
  };
  auto eval = [&](const double *x, double *g) { return rosen(x, g, n); };
  BfgsResult ra = lbfgs_solve(n, xa.data(), eval, check);
  bool ok = ra.status == kConverged;
  for (int lose : {0, 5}) {
    std::vector<double> xb;
    int it = 0, ev = 0;
    const int st = served(n, tol, max_it, lose, xb, it, ev);
    const bool same = xb == xa && it == ra.iterations && ev == ra.evaluations && st == ra.status;
    std::printf("loop: it=%d ev=%d st=%d  served(lose every %d): it=%d ev=%d st=%d  same=%d\n",
                ra.iterations, ra.evaluations, ra.status, lose, it, ev, st, (int)same);
    ok = ok && same;
  }
  // diagonal initial inverse Hessian: the driver with h0 / nu_j reproduces
  // lbfgs_solve with h0[j] = h0 / nu_j; on a separable quadratic whose Hessian
  // is diag(nu_j / h0) the first step (trial step 1) lands on the minimiser
  {
    std::vector<double> nu(n), h(n);
    for (int64_t i = 0; i < n; ++i) {
      nu[i] = 0.5 + 0.37 * i;
      h[i] = 0.25 / nu[i];
    }
    std::vector<double> xp(n, -1.0);
    xp[1] = 1.2;
    BfgsResult rp = lbfgs_solve(n, xp.data(), eval, check, 5, 0.0, h.data());
    std::vector<double> xs;
    int it = 0, ev = 0;
    const int st = served(n, tol, max_it, 0, xs, it, ev, 0.0, 0.25, &nu);
    const bool same = xs == xp && it == rp.iterations && ev == rp.evaluations && st == rp.status;
    auto quad = [&](const double *x, double *g) {
      double f = 0;
      for (int64_t i = 0; i < n; ++i) {
        const double d = x[i] - (0.1 * i - 0.3);
        g[i] = nu[i] / 0.25 * d;
        f += 0.5 * nu[i] / 0.25 * d * d;
      }
      return f;
    };
    std::vector<double> xq(n, 2.0);
    BfgsResult rq = lbfgs_solve(n, xq.data(), quad, check, 5, 0.0, h.data());
    double err = 0;
    for (int64_t i = 0; i < n; ++i) err = std::fmax(err, std::fabs(xq[i] - (0.1 * i - 0.3)));
    std::printf("diagonal H0: loop it=%d ev=%d st=%d  served it=%d ev=%d st=%d  same=%d; "
                "quadratic it=%d ev=%d err=%.1e\n", rp.iterations, rp.evaluations, rp.status,
                it, ev, st, (int)same, rq.iterations, rq.evaluations, err);
    ok = ok && same && rp.status == kConverged && rq.status == kConverged &&
         rq.iterations == 1 && rq.evaluations == 2 && err <= 1e-15;
  }
  // noise-aware line search on Rosenbrock + 1000 with a deterministic
  // rounding-like perturbation of 1e-7 on F (gradient exact): the reference
  // search runs out of evaluations once the decrease is below the noise; the
  // noise-aware one converges; loop form and driver take identical decisions
  {
    auto noisy = [&](const double *x, double *g) {
      double h = 0;
      for (int64_t i = 0; i < n; ++i) h += x[i] * (12.9898 + 7.233 * i);
      const double r = std::sin(h) * 43758.5453;
      return rosen(x, g, n) + 1000.0 + 1e-7 * (r - std::floor(r) - 0.5);
    };
    const double tight = 1e-7;
    auto check_t = [&](int step, const double *g, double &cv) {
      cv = 0;
      for (int64_t i = 0; i < n; ++i) cv += std::fabs(g[i]);
      if (cv <= tight) return (int)kSuccess;
      if ((uint32_t)step >= max_it || std::isnan(cv)) return (int)kFailure;
      return (int)kIterate;
    };
    std::vector<double> xr(n, -1.0);
    xr[1] = 1.2;
    BfgsResult rr = lbfgs_solve(n, xr.data(), noisy, check_t);
    const double nr = 1e-9;
    std::vector<double> xn(n, -1.0);
    xn[1] = 1.2;
    BfgsResult rn = lbfgs_solve(n, xn.data(), noisy, check_t, 5, nr);
    // the driver with the same noisy objective
    Emu e(n);
    for (auto &v : e.x) v = -1.0;
    e.x[1] = 1.2;
    SolveDriver d;
    d.init(0, tight, max_it, nr);
    int st = -1, raw = 0;
    while (st < 0) {
      const SlotIn c = d.next_in();
      SlotOut o{};
      e.action(c, o);
      e.evaluate(o, tight);
      o.F = noisy(e.x.data(), e.g.data());
      ++raw;
      st = d.on_result(o, raw % 4 == 0);
    }
    const bool same = e.x == xn && d.iterations() == rn.iterations &&
                      d.evaluations() == rn.evaluations && st == rn.status;
    std::printf("noisy F: reference st=%d it=%d; noise-aware loop st=%d it=%d ev=%d  driver st=%d "
                "it=%d ev=%d  same=%d\n", rr.status, rr.iterations, rn.status, rn.iterations,
                rn.evaluations, st, d.iterations(), d.evaluations(), (int)same);
    ok = ok && same && rr.status == kLineSearchFail && rn.status == kConverged;
  }
  // precision escalation: "low precision" evaluations carry noise on F (1e-7)
  // and on every gradient component (1e-6); the driver must ask for exactly
  // one switch and then stop on an exact evaluation within the tolerance
  {
    const double tight = 1e-7;
    bool low = true;
    auto evalp = [&](const double *x, double *g) {
      double f = rosen(x, g, n) + 1000.0;
      if (low) {
        double h = 0;
        for (int64_t i = 0; i < n; ++i) h += x[i] * (12.9898 + 7.233 * i);
        double r = std::sin(h) * 43758.5453;
        f += 1e-7 * (r - std::floor(r) - 0.5);
        for (int64_t i = 0; i < n; ++i) {
          r = std::sin(h + 1.7 * (i + 1)) * 24634.6345;
          g[i] += 1e-6 * (r - std::floor(r) - 0.5);
        }
      }
      return f;
    };
    Emu e(n);
    for (auto &v : e.x) v = -1.0;
    e.x[1] = 1.2;
    SolveDriver d;
    d.init(0, tight, max_it, 1e-9, 10.0, 1e-15);
    int st = -1, switches = 0, exact_evals = 0;
    double last_l1 = 0;
    while (st < 0) {
      const SlotIn c = d.next_in();
      SlotOut o{};
      e.action(c, o);
      e.evaluate(o, tight);
      o.F = evalp(e.x.data(), e.g.data());
      o.gp = e.dot(e.g, e.p);
      o.l1 = 0;
      for (int64_t i = 0; i < n; ++i) o.l1 += std::fabs(e.g[i]);
      if (!low) ++exact_evals;
      last_l1 = o.l1;
      st = d.on_result(o, false);
      if (st == SolveDriver::kSwitch) {
        ++switches;
        low = false;
        st = -1;
      }
    }
    std::printf("escalation: st=%d it=%d ev=%d switches=%d exact evals=%d final l1=%.3e\n", st,
                d.iterations(), d.evaluations(), switches, exact_evals, last_l1);
    ok = ok && st == kConverged && switches == 1 && !low && last_l1 <= tight && exact_evals >= 2;
  }
  // iteration cap: both stop with MAX_ITER after the same number of steps
  {
    std::vector<double> xc(n, -1.0);
    xc[1] = 1.2;
    auto check7 = [&](int step, const double *g, double &cv) {
      cv = 0;
      for (int64_t i = 0; i < n; ++i) cv += std::fabs(g[i]);
      if (cv <= tol) return (int)kSuccess;
      if (step >= 7 || std::isnan(cv)) return (int)kFailure;
      return (int)kIterate;
    };
    BfgsResult rc = lbfgs_solve(n, xc.data(), eval, check7);
    std::vector<double> xb;
    int it = 0, ev = 0;
    const int st = served(n, tol, 7, 0, xb, it, ev);
    const bool same = xb == xc && it == rc.iterations && ev == rc.evaluations && st == rc.status;
    std::printf("cap 7: loop it=%d ev=%d st=%d  served it=%d ev=%d st=%d  same=%d\n", rc.iterations,
                rc.evaluations, rc.status, it, ev, st, (int)same);
    ok = ok && same && st == kMaxIter;
  }
  return ok ? 0 : 1;
}
