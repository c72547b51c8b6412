// Drives LineSearchSM with host vectors and compares every iterate with the
// loop form lbfgs_solve on Rosenbrock and a quadratic (bitwise).
#include <cstdio>
#include <vector>
#include "../semidiscreteot_amd/csrc/sdot_bfgs.hpp"
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

int main() {
  const int64_t n = 6;
  std::vector<double> xa(n, -1.0), xb(n, -1.0);
  xa[1] = xb[1] = 1.2;
  std::vector<std::vector<double>> trace_a;
  auto eval_a = [&](const double *x, double *g) { trace_a.emplace_back(x, x + n); return rosen(x, g, n); };
  auto check = [&](int step, const double *g, double &cv) {
    cv = 0; for (int64_t i = 0; i < n; ++i) cv += g[i] * g[i];
    cv = std::sqrt(cv);
    if (cv <= 1e-10) return (int)kSuccess;
    if (step >= 500) return (int)kFailure;
    return (int)kIterate;
  };
  BfgsResult ra = lbfgs_solve(n, xa.data(), eval_a, check);

  // lock-step form
  std::vector<std::vector<double>> trace_b;
  const int h = 5;
  std::vector<double> g(n), p(n), x0(n), sk(n), yk(n), rho(h), coef(h);
  std::vector<std::vector<double>> S(h, std::vector<double>(n)), Y(h, std::vector<double>(n));
  int count = 0, newest = -1, k = 0, evals = 0;
  auto slot = [&](int i) { return ((newest - i) % h + h) % h; };
  trace_b.emplace_back(xb);
  double f = rosen(xb.data(), g.data(), n); ++evals;
  double cv; int st = check(0, g.data(), cv);
  bool first = true; double f_prev = 0;
  LineSearchSM ls;
  int status = kMaxIter;
  while (st == kIterate) {
    p = g;
    for (int i = 0; i < count; ++i) { int s = slot(i); coef[i] = rho[s] * bfgs_dot(S[s].data(), p.data(), n);
      for (int64_t t = 0; t < n; ++t) p[t] += -coef[i] * Y[s][t]; }
    for (int i = count - 1; i >= 0; --i) { int s = slot(i); double c2 = rho[s] * bfgs_dot(Y[s].data(), p.data(), n);
      double c = coef[i] - c2; for (int64_t t = 0; t < n; ++t) p[t] += c * S[s][t]; }
    for (auto &v : p) v = -v;
    sk = xb; yk = g; x0 = xb;
    const double g0 = bfgs_dot(g.data(), p.data(), n);
    const double df = std::fmax(f_prev - f, 100.0 * DBL_EPSILON);
    const double a1 = first ? 1.0 : std::fmin(1.0, -1.01 * 2.0 * df / g0);
    f_prev = f;
    ls.start(f, g0, a1);
    int r;
    do {
      for (int64_t t = 0; t < n; ++t) xb[t] = x0[t] + ls.trial() * p[t];
      trace_b.emplace_back(xb);
      f = rosen(xb.data(), g.data(), n); ++evals;
      r = ls.feed(f, bfgs_dot(g.data(), p.data(), n));
    } while (r == LineSearchSM::kEval);
    if (r == LineSearchSM::kFail) { ++k; status = kLineSearchFail; break; }
    first = false;
    for (int64_t t = 0; t < n; ++t) { sk[t] = xb[t] - sk[t]; yk[t] = g[t] - yk[t]; }
    ++k;
    st = check(k, g.data(), cv);
    if (st != kIterate) { status = st == kSuccess ? kConverged : kMaxIter; break; }
    double curv = bfgs_dot(sk.data(), yk.data(), n);
    if (curv > 0) { newest = (newest + 1) % h; S[newest] = sk; Y[newest] = yk; rho[newest] = 1 / curv; if (count < h) ++count; }
  }
  bool same = trace_a == trace_b && ra.iterations == k && ra.evaluations == evals && ra.status == status;
  std::printf("loop: it=%d ev=%d st=%d  lockstep: it=%d ev=%d st=%d  same=%d x0=%.17g\n",
              ra.iterations, ra.evaluations, ra.status, k, evals, status, (int)same, xb[0]);
  return same && status == kConverged ? 0 : 1;
}
