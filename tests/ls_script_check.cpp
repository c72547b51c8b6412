// Runs the library's L-BFGS / line search (lbfgs_solve of sdot_bfgs.hpp, the
// FP64 host driver of sdot_solve) on a scripted objective: the k-th
// evaluation returns the k-th (f, g) of the script whatever x it is asked
// at, and every x asked for is printed.  The calling test compares the
// printed trial points with hand-derived ones.
//
// stdin: n max_iterations tol_l2
//        x0[0..n)
//        n_entries, then n_entries lines "f g[0..n)"
// stdout: one line per evaluation "x[0..n)", then "status iterations evaluations"
#include <cstdio>
#include <vector>

#include "../semidiscreteot_amd/csrc/sdot_bfgs.hpp"

int main() {
  long n = 0, max_it = 0, m = 0;
  double tol = 0;
  if (std::scanf("%ld %ld %lf", &n, &max_it, &tol) != 3 || n <= 0) return 2;
  std::vector<double> x(n);
  for (long i = 0; i < n; ++i)
    if (std::scanf("%lf", &x[i]) != 1) return 2;
  if (std::scanf("%ld", &m) != 1 || m <= 0) return 2;
  std::vector<double> fs(m), gs(m * n);
  for (long k = 0; k < m; ++k) {
    if (std::scanf("%lf", &fs[k]) != 1) return 2;
    for (long i = 0; i < n; ++i)
      if (std::scanf("%lf", &gs[k * n + i]) != 1) return 2;
  }
  long used = 0;
  auto eval = [&](const double *xv, double *g) {
    for (long i = 0; i < n; ++i) std::printf("%s%.17g", i ? " " : "", xv[i]);
    std::printf("\n");
    const long k = used < m ? used : m - 1;  // past the script: repeat its last entry
    ++used;
    for (long i = 0; i < n; ++i) g[i] = gs[k * n + i];
    return fs[k];
  };
  auto check = [&](int step, const double *g, double &cv) {
    cv = std::sqrt(sdot::bfgs_dot(g, g, n));
    if (cv <= tol) return (int)sdot::kSuccess;
    if (step >= max_it) return (int)sdot::kFailure;
    return (int)sdot::kIterate;
  };
  const sdot::BfgsResult r = sdot::lbfgs_solve(n, x.data(), eval, check);
  std::printf("%d %d %d\n", r.status, r.iterations, r.evaluations);
  return 0;
}
