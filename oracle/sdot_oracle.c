/*
 * sdot_oracle.c -- FP64 CPU restatement of SemiDiscreteOT's regularized
 * semi-discrete dual evaluation, its truncation thresholds and its BFGS
 * driver.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the "port" CPU
 * baseline.  The product library never links or loads this file.
 *
 * What pins each function (a function with no independent check would say
 * "parity unpinned"; every function here has at least one):
 *   orc_gauss_legendre01      textbook nodes/weights (2- and 3-point closed
 *                             forms), exact integration of x^(2n-1).
 *   orc_structured_source     closed-form integrals of polynomials over the
 *                             box; sum of JxW = volume.
 *   orc_simplex_source        Keast 4-point rule closed form; tet volumes.
 *   orc_cells_source          affine Q1 boxes equal orc_structured_source;
 *                             volumes and moments of trilinear hexes with
 *                             a raised vertex (closed forms); areas of tilted
 *                             and of bilinear (saddle) surface quads in 3D
 *                             (closed form, 50-digit quadrature); tets equal
 *                             orc_simplex_source; surface triangle areas.
 *   orc_evaluate(_R)          N=1 closed form, quadrant-symmetric g(0)=0,
 *                             sum(g)=mass-1, F(psi+c)=F(psi), g = FD of F,
 *                             50-digit brute force on tiny inputs
 *                             (tests/golden), plain vs LSE mode agreement.
 *   orc_candidate_stats       brute-force Python scan with the same rule.
 *   orc_distance_threshold    closed-form evaluation of the three formulas.
 *   orc_covering_radius       numpy brute force; one central target (closed
 *                             form sqrt(3)(1/2 - 1/2n)); targets at the centers.
 *   orc_discarded_fraction    numpy brute force; R = 10 and R < 0 limits; two
 *                             symmetric targets (closed-form logistic share).
 *   orc_hessian               FD of the gradient, H*1 = 0, symmetry, PSD.
 *   orc_bfgs_generic/solve    hand-derived line-search traces on scripted
 *                             objectives (acceptance, growth 9, safeguards
 *                             0.1/0.5, cubic / quadratic / bisection, first
 *                             step 2.02, eta 0.9, mu 1e-3, 20-evaluation cap)
 *                             and the two-loop direction against the explicit
 *                             BFGS matrix update (tests/test_line_search_pins.py);
 *                             quadratic (-> A^-1 b) and Rosenbrock minima,
 *                             converged SOT gradients below tolerance.
 *   orc_barycenters_R         one target: the mean of the source (sum w x);
 *                             a shift of every point shifts the barycenters;
 *                             sum_j nu_j b_j = sum_q w_q x_q when the plan's
 *                             marginal is exact; first moments vs gradient.
 *   orc_conditional_density   partition of unity: sum_i nu_i c_i(x) = rho(x)
 *                             over all targets; one target: rho; 50-digit
 *                             values (tests/golden).
 *   orc_transport_map         modal = nearest target when psi = eps log nu
 *                             terms cancel (Voronoi); barycentric of one
 *                             target = that target; eps -> 0 barycentric =
 *                             modal; 50-digit values (tests/golden).
 *   orc_softmax_refinement    identity hierarchy (fine = coarse, one child
 *                             each): psi_f = psi_c - lambda log(1 + g_c/nu_c)
 *                             with g from orc_evaluate_R at the same psi;
 *                             a child listed under two parents gets
 *                             -lambda log 2 more; 50-digit brute force with
 *                             truncation on tiny inputs (tests/golden).
 * Only the last bits of summation order (R-tree traversal order, WorkStream
 * copier order, deal.II's blocked dot products) are not reproduced; they are
 * far below the 1e-10 parity bar.
 */
#include "sdot_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int orc_threads(const orc_params_t *p) {
#ifdef _OPENMP
  int n = (p && p->n_threads > 0) ? p->n_threads : omp_get_max_threads();
  return n < 1 ? 1 : n;
#else
  (void)p;
  return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Quadrature (deal.II QGauss / QGaussSimplex, not vendored in the reference;
 * chosen by SotSolver.cc:243-252 and utils.h:818-840).                      */
/* ------------------------------------------------------------------------ */

/* Gauss-Legendre nodes/weights mapped to [0,1], ascending. Newton iteration
 * on the Legendre recurrence. */
void orc_gauss_legendre01(int n, double *pts, double *wts) {
  for (int i = 0; i < n; ++i) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5));
    double dp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = 0.0;
      for (int k = 1; k <= n; ++k) {
        double p2 = p1;
        p1 = p0;
        p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
      }
      dp = n * (z * p0 - p1) / (z * z - 1.0);
      double dz = p0 / dp;
      z -= dz;
      if (fabs(dz) < 1e-17) break;
    }
    {
      double p0 = 1.0, p1 = 0.0;
      for (int k = 1; k <= n; ++k) {
        double p2 = p1;
        p1 = p0;
        p0 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p2) / k;
      }
      dp = n * (z * p0 - p1) / (z * z - 1.0);
    }
    /* node z in (-1,1) descending with i; map to [0,1] ascending */
    int idx = n - 1 - i;
    pts[idx] = 0.5 * (1.0 - z) ;
    wts[idx] = 1.0 / ((1.0 - z * z) * dp * dp);
  }
  /* ascending order: pts computed as (1-z)/2 with z descending -> ascending
   * already when indexed by i; re-sort defensively */
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j)
      if (pts[j] < pts[i]) {
        double a = pts[i]; pts[i] = pts[j]; pts[j] = a;
        a = wts[i]; wts[i] = wts[j]; wts[j] = a;
      }
}

/* Structured Q1 mesh of the box [lo,hi] with n_sub cells per direction
 * (deal.II GridGenerator::subdivided_hyper_rectangle; cells and vertices in
 * lexicographic order, x fastest), QGauss(qorder) tensor rule, Q1 density
 * interpolation (SotSolver.cc:788-790 via FEValues::get_function_values),
 * cell center = vertex average (deal.II TriaAccessor::center). */
int orc_structured_source(int dim, const int *n_sub, const double *lo,
                          const double *hi, int qorder,
                          const double *nodal_rho, double *x, double *rho,
                          double *jxw, double *center) {
  if (dim != 2 && dim != 3) return -1;
  if (qorder < 1 || qorder > 10) return -1;
  double gp[10], gw[10];
  orc_gauss_legendre01(qorder, gp, gw);
  int ns[3] = {n_sub[0], n_sub[1], dim == 3 ? n_sub[2] : 1};
  double h[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) h[d] = (hi[d] - lo[d]) / ns[d];
  const int nv = 1 << dim;
  int nq = 1;
  for (int d = 0; d < dim; ++d) nq *= qorder;
  const int64_t nvx = ns[0] + 1, nvy = ns[1] + 1;
  const int64_t n_cells = (int64_t)ns[0] * ns[1] * ns[2];
  double detJ = 1.0;
  for (int d = 0; d < dim; ++d) detJ *= h[d];

#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < n_cells; ++c) {
    int64_t ci = c % ns[0];
    int64_t cj = (c / ns[0]) % ns[1];
    int64_t ck = c / ((int64_t)ns[0] * ns[1]);
    double vx[8][3];
    double vr[8];
    for (int v = 0; v < nv; ++v) {
      int bx = v & 1, by = (v >> 1) & 1, bz = (v >> 2) & 1;
      int64_t ii = ci + bx, jj = cj + by, kk = ck + bz;
      vx[v][0] = lo[0] + ii * h[0];
      vx[v][1] = lo[1] + jj * h[1];
      vx[v][2] = dim == 3 ? lo[2] + kk * h[2] : 0.0;
      int64_t vid = ii + nvx * (jj + nvy * (dim == 3 ? kk : 0));
      vr[v] = nodal_rho ? nodal_rho[vid] : 1.0;
    }
    double cc[3] = {0, 0, 0};
    for (int v = 0; v < nv; ++v)
      for (int d = 0; d < dim; ++d) cc[d] += vx[v][d];
    for (int d = 0; d < dim; ++d) center[c * dim + d] = cc[d] / nv;
    for (int q = 0; q < nq; ++q) {
      int qi = q % qorder, qj = (q / qorder) % qorder, qk = q / (qorder * qorder);
      double xi[3] = {gp[qi], gp[qj], dim == 3 ? gp[qk] : 0.0};
      double wq = gw[qi] * gw[qj] * (dim == 3 ? gw[qk] : 1.0);
      double xp[3] = {0, 0, 0}, r = 0.0;
      for (int v = 0; v < nv; ++v) {
        double phi = 1.0;
        for (int d = 0; d < dim; ++d) phi *= ((v >> d) & 1) ? xi[d] : 1.0 - xi[d];
        for (int d = 0; d < dim; ++d) xp[d] += vx[v][d] * phi;
        r += vr[v] * phi;
      }
      int64_t g = c * nq + q;
      for (int d = 0; d < dim; ++d) x[g * dim + d] = xp[d];
      rho[g] = r;
      jxw[g] = detJ * wq;
    }
  }
  return 0;
}

/* P1 simplices (deal.II FE_SimplexP(1) + MappingFE, QGaussSimplex(qorder)),
 * utils.h:776-840.  qorder 1 -> centroid; 2 -> 3-point (triangles) or the
 * 4-point rule with a=(5-sqrt5)/20, b=(5+3sqrt5)/20 (tets). */
int orc_simplex_source(int dim, int64_t n_vertices, const double *vertices,
                       int64_t n_cells, const int64_t *cells, int qorder,
                       const double *nodal_rho, double *x, double *rho,
                       double *jxw, double *center) {
  (void)n_vertices;
  if (dim != 2 && dim != 3) return -1;
  double qp[4][3];
  double qw[4];
  int nq;
  const int nv = dim + 1;
  if (qorder == 1) {
    nq = 1;
    for (int d = 0; d < 3; ++d) qp[0][d] = d < dim ? 1.0 / (dim + 1) : 0.0;
    qw[0] = dim == 2 ? 0.5 : 1.0 / 6.0;
  } else if (qorder == 2 && dim == 2) {
    nq = 3;
    const double a = 1.0 / 6.0, b = 2.0 / 3.0;
    double p[3][2] = {{a, a}, {b, a}, {a, b}};
    for (int q = 0; q < 3; ++q) {
      qp[q][0] = p[q][0];
      qp[q][1] = p[q][1];
      qp[q][2] = 0;
      qw[q] = 1.0 / 6.0;
    }
  } else if (qorder == 2 && dim == 3) {
    nq = 4;
    const double a = (5.0 - sqrt(5.0)) / 20.0, b = (5.0 + 3.0 * sqrt(5.0)) / 20.0;
    double p[4][3] = {{a, a, a}, {b, a, a}, {a, b, a}, {a, a, b}};
    for (int q = 0; q < 4; ++q) {
      for (int d = 0; d < 3; ++d) qp[q][d] = p[q][d];
      qw[q] = 1.0 / 24.0;
    }
  } else {
    return -1;
  }
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < n_cells; ++c) {
    double v[4][3];
    double vr[4];
    for (int i = 0; i < nv; ++i) {
      int64_t vid = cells[c * nv + i];
      for (int d = 0; d < 3; ++d) v[i][d] = d < dim ? vertices[vid * dim + d] : 0.0;
      vr[i] = nodal_rho ? nodal_rho[vid] : 1.0;
    }
    double cc[3] = {0, 0, 0};
    for (int i = 0; i < nv; ++i)
      for (int d = 0; d < dim; ++d) cc[d] += v[i][d];
    for (int d = 0; d < dim; ++d) center[c * dim + d] = cc[d] / nv;
    double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int d = 0; d < dim; ++d)
      for (int k = 0; k < dim; ++k) J[d][k] = v[k + 1][d] - v[0][d];
    double det;
    if (dim == 2)
      det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    else
      det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
            J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
            J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    det = fabs(det);
    for (int q = 0; q < nq; ++q) {
      double phi[4];
      phi[0] = 1.0;
      for (int k = 0; k < dim; ++k) {
        phi[k + 1] = qp[q][k];
        phi[0] -= qp[q][k];
      }
      double xp[3] = {0, 0, 0}, r = 0.0;
      for (int i = 0; i < nv; ++i) {
        for (int d = 0; d < dim; ++d) xp[d] += v[i][d] * phi[i];
        r += vr[i] * phi[i];
      }
      int64_t g = c * nq + q;
      for (int d = 0; d < dim; ++d) x[g * dim + d] = xp[d];
      rho[g] = r;
      jxw[g] = det * qw[q];
    }
  }
  return nq;
}

/* L1 norm used by normalize_density (SemiDiscreteOT.cc:422-449): the caller
 * divides the nodal density by it. */
/* General cells as GridIn::read_msh / read_vtk meshes reach setup_source
 * (MeshManager.templates.h:80-140), with the FEValues values of
 * SotSolver.cc:788-790: q1 != 0: FE_Q(1) / MappingQ1 / QGauss(qorder) on
 * 2^cdim vertices in lexicographic order; q1 == 0: FE_SimplexP(1) /
 * QGaussSimplex(qorder) on cdim + 1 vertices.  Cells of dimension cdim in a
 * space of dimension sdim (surfaces in 3D: SotSolver<2,3>).
 * The Jacobian of the multilinear map is taken column by column as the
 * difference of the two opposite faces' interpolants (exact for a map linear
 * in each reference coordinate); the measure is |det J| for cdim == sdim and
 * the cross-product length |J_0 x J_1| for a surface in 3D.  Returns nq. */
static void q1_point(int cdim, int sdim, double v[8][3], const double *xi, double *out) {
  for (int d = 0; d < sdim; ++d) out[d] = 0.0;
  for (int k = 0; k < (1 << cdim); ++k) {
    double phi = 1.0;
    for (int a = 0; a < cdim; ++a) phi *= ((k >> a) & 1) ? xi[a] : 1.0 - xi[a];
    for (int d = 0; d < sdim; ++d) out[d] += phi * v[k][d];
  }
}

int orc_cells_source(int q1, int cdim, int sdim, int64_t n_vertices,
                     const double *vertices, int64_t n_cells, const int64_t *cells,
                     int qorder, const double *nodal_rho, double *x, double *rho,
                     double *jxw, double *center) {
  (void)n_vertices;
  if (cdim < 2 || cdim > 3 || sdim < cdim || sdim > 3) return -1;
  double qp[1000][3], qw[1000];
  int nq = 0;
  if (q1) {
    if (qorder < 1 || qorder > 10) return -1;
    double gp[10], gw[10];
    orc_gauss_legendre01(qorder, gp, gw);
    const int nz = cdim == 3 ? qorder : 1;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < qorder; ++j)
        for (int i = 0; i < qorder; ++i) {
          qp[nq][0] = gp[i];
          qp[nq][1] = gp[j];
          qp[nq][2] = cdim == 3 ? gp[k] : 0.0;
          qw[nq] = gw[i] * gw[j] * (cdim == 3 ? gw[k] : 1.0);
          ++nq;
        }
  } else {
    /* the QGaussSimplex rules of orc_simplex_source */
    if (qorder == 1) {
      nq = 1;
      for (int d = 0; d < 3; ++d) qp[0][d] = d < cdim ? 1.0 / (cdim + 1) : 0.0;
      qw[0] = cdim == 2 ? 0.5 : 1.0 / 6.0;
    } else if (qorder == 2 && cdim == 2) {
      const double a = 1.0 / 6.0, b = 2.0 / 3.0;
      const double p[3][2] = {{a, a}, {b, a}, {a, b}};
      nq = 3;
      for (int q = 0; q < 3; ++q) {
        qp[q][0] = p[q][0];
        qp[q][1] = p[q][1];
        qp[q][2] = 0.0;
        qw[q] = 1.0 / 6.0;
      }
    } else if (qorder == 2 && cdim == 3) {
      const double a = (5.0 - sqrt(5.0)) / 20.0, b = (5.0 + 3.0 * sqrt(5.0)) / 20.0;
      const double p[4][3] = {{a, a, a}, {b, a, a}, {a, b, a}, {a, a, b}};
      nq = 4;
      for (int q = 0; q < 4; ++q) {
        for (int d = 0; d < 3; ++d) qp[q][d] = p[q][d];
        qw[q] = 1.0 / 24.0;
      }
    } else {
      return -1;
    }
  }
  const int nv = q1 ? 1 << cdim : cdim + 1;
#pragma omp parallel for schedule(static)
  for (int64_t c = 0; c < n_cells; ++c) {
    double v[8][3] = {{0}}, vr[8], cc[3] = {0, 0, 0};
    for (int i = 0; i < nv; ++i) {
      const int64_t vid = cells[c * nv + i];
      for (int d = 0; d < sdim; ++d) v[i][d] = vertices[vid * sdim + d];
      vr[i] = nodal_rho ? nodal_rho[vid] : 1.0;
      for (int d = 0; d < sdim; ++d) cc[d] += v[i][d];
    }
    for (int d = 0; d < sdim; ++d) center[c * sdim + d] = cc[d] / nv;
    for (int q = 0; q < nq; ++q) {
      double xp[3] = {0, 0, 0}, J[3][3] = {{0}}, r = 0.0;
      if (q1) {
        q1_point(cdim, sdim, v, qp[q], xp);
        for (int a = 0; a < cdim; ++a) {
          double lo_xi[3], hi_xi[3], plo[3], phi_[3];
          for (int b = 0; b < 3; ++b) lo_xi[b] = hi_xi[b] = qp[q][b];
          lo_xi[a] = 0.0;
          hi_xi[a] = 1.0;
          q1_point(cdim, sdim, v, lo_xi, plo);
          q1_point(cdim, sdim, v, hi_xi, phi_);
          for (int d = 0; d < sdim; ++d) J[d][a] = phi_[d] - plo[d];
        }
        for (int k = 0; k < nv; ++k) {
          double phi = 1.0;
          for (int a = 0; a < cdim; ++a) phi *= ((k >> a) & 1) ? qp[q][a] : 1.0 - qp[q][a];
          r += phi * vr[k];
        }
      } else {
        double lam[4];
        lam[0] = 1.0;
        for (int a = 0; a < cdim; ++a) {
          lam[a + 1] = qp[q][a];
          lam[0] -= qp[q][a];
        }
        for (int k = 0; k < nv; ++k) {
          for (int d = 0; d < sdim; ++d) xp[d] += lam[k] * v[k][d];
          r += lam[k] * vr[k];
        }
        for (int a = 0; a < cdim; ++a)
          for (int d = 0; d < sdim; ++d) J[d][a] = v[a + 1][d] - v[0][d];
      }
      double meas;
      if (cdim == sdim) {
        meas = cdim == 2 ? J[0][0] * J[1][1] - J[0][1] * J[1][0]
                         : J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                               J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                               J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        meas = fabs(meas);
      } else {
        const double cx = J[1][0] * J[2][1] - J[2][0] * J[1][1];
        const double cy = J[2][0] * J[0][1] - J[0][0] * J[2][1];
        const double cz = J[0][0] * J[1][1] - J[1][0] * J[0][1];
        meas = sqrt(cx * cx + cy * cy + cz * cz);
      }
      const int64_t g = c * nq + q;
      for (int d = 0; d < sdim; ++d) x[g * sdim + d] = xp[d];
      rho[g] = r;
      jxw[g] = meas * qw[q];
    }
  }
  return nq;
}

double orc_normalize_density(int64_t n_q, const double *rho, const double *jxw) {
  double s = 0.0;
  for (int64_t i = 0; i < n_q; ++i) s += fabs(rho[i]) * jxw[i];
  return s;
}

/* ------------------------------------------------------------------------ */
/* Distances and thresholds                                                  */
/* ------------------------------------------------------------------------ */

/* euclidean_distance (Distance.h:23-27): deal.II (a-b).norm() =
 * sqrt(((0 + d0*d0) + d1*d1) + d2*d2), no FMA contraction. */
static inline double orc_dist(const double *a, const double *b, int dim) {
  double s = 0.0;
  for (int d = 0; d < dim; ++d) {
    double t = a[d] - b[d];
    s += t * t;
  }
  return sqrt(s);
}

static double vec_max(const double *v, int64_t n) {
  double m = v[0];
  for (int64_t i = 1; i < n; ++i) m = v[i] > m ? v[i] : m;
  return m;
}
static double vec_min(const double *v, int64_t n) {
  double m = v[0];
  for (int64_t i = 1; i < n; ++i) m = v[i] < m ? v[i] : m;
  return m;
}

/* compute_covering_radius (SotSolver.cc:431-473) without the 1.1 factor that
 * solve() applies (SotSolver.cc:132): max over cells of the distance from the
 * cell center to its nearest target (brute force instead of the R-tree). */
double orc_covering_radius(const orc_source_t *s, const orc_target_t *t) {
  double best = 0.0;
  const int dim = s->dim;
#pragma omp parallel for schedule(static) reduction(max : best)
  for (int64_t c = 0; c < s->n_cells; ++c) {
    const double *cc = s->center + c * dim;
    double md = INFINITY;
    for (int64_t j = 0; j < t->n; ++j) {
      double d = orc_dist(t->y + j * dim, cc, dim);
      if (d < md) md = d;
    }
    if (md > best) best = md;
  }
  return best;
}

/* compute_integral_radius_bound (SotSolver.cc:475-507) */
static double orc_integral_bound(const double *psi, int64_t n, double eps,
                                 double tol, double C, double Fprev) {
  double max_potential = vec_max(psi, n);
  double absF = fabs(Fprev);
  if (absF < 1e-10) absF = 1e-10;
  double arg = (eps * C) / (tol * absF);
  double r2 = 2.0 * max_potential + 2.0 * eps * log(arg);
  return sqrt(r2 > 0.0 ? r2 : 0.0);
}

/* compute_geometric_radius_bound (SotSolver.cc:509-551) */
static double orc_geometric_bound(const orc_source_t *s, const orc_target_t *t,
                                  const orc_state_t *st, const double *psi,
                                  double eps, double tol) {
  double r0 = st->covering_radius > 0.0 ? st->covering_radius : orc_covering_radius(s, t);
  double mn = vec_min(psi, t->n), mx = vec_max(psi, t->n);
  double range = mx - mn;
  double min_density = st->min_target_density > 0.0 ? st->min_target_density : vec_min(t->nu, t->n);
  double fv = fabs(st->global_functional);
  if (fv < 1e-10) fv = 1e-10;
  double log_term = log(eps / (min_density * tol * fv));
  double r2 = r0 * r0 + 2.0 * range + 2.0 * eps * log_term;
  if (r2 < r0 * r0) r2 = r0 * r0;
  return sqrt(r2);
}

/* compute_distance_threshold (SotSolver.cc:379-429); reads the state left by
 * the previous evaluation (C_global, global_functional). */
double orc_distance_threshold(const orc_params_t *p, const orc_state_t *st,
                              const orc_source_t *s, const orc_target_t *t,
                              const double *psi) {
  const double eps = p->epsilon;
  double R;
  if (p->threshold_type == ORC_THRESH_INTEGRAL) {
    R = orc_integral_bound(psi, t->n, eps, p->tau, st->C_global, st->global_functional);
    R = R > st->covering_radius ? R : st->covering_radius;
  } else if (p->threshold_type == ORC_THRESH_GEOMETRIC) {
    R = orc_geometric_bound(s, t, st, psi, eps, p->tau);
  } else {
    double max_potential = vec_max(psi, t->n);
    double mind = st->min_target_density > 0.0 ? st->min_target_density : vec_min(t->nu, t->n);
    if (mind <= 0 || p->tau <= 0) {
      R = DBL_MAX;
    } else {
      double r2 = -2.0 * eps * log(p->tau / mind) + 2.0 * max_potential;
      R = sqrt(r2 > 0.0 ? r2 : 0.0);
    }
  }
  return R;
}

/* find_nearest_target_points (SotSolver.cc:553-570): every target whose
 * distance to the query is <= R, in index order. */
static int64_t orc_candidates(const orc_target_t *t, const double *c, int dim,
                              double R, int64_t *out) {
  int64_t n = 0;
  for (int64_t j = 0; j < t->n; ++j)
    if (orc_dist(t->y + j * dim, c, dim) <= R) out[n++] = j;
  return n;
}

int orc_candidate_stats(const orc_source_t *s, const orc_target_t *t, double R,
                        int64_t *count, uint64_t *hash) {
  const int dim = s->dim;
#pragma omp parallel
  {
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (t->n > 0 ? t->n : 1));
#pragma omp for schedule(static)
    for (int64_t c = 0; c < s->n_cells; ++c) {
      int64_t n = orc_candidates(t, s->center + c * dim, dim, R, idx);
      uint64_t h = 0;
      for (int64_t k = 0; k < n; ++k) {
        uint64_t v = (uint64_t)idx[k] + 1u;
        v *= 0x9E3779B97F4A7C15ull;
        v ^= v >> 29;
        h += v;
      }
      count[c] = n;
      hash[c] = h;
    }
    free(idx);
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* evaluate_functional (SotSolver.cc:217-377) with local_assemble             */
/* (SotSolver.cc:769-869) and the function_call lambda (SotSolver.cc:267-296) */
/* ------------------------------------------------------------------------ */

int orc_evaluate_R(const orc_source_t *s, const orc_target_t *t,
                   const orc_params_t *p, double R, const double *psi,
                   double *grad, double *F, double *C) {
  const int dim = s->dim;
  const int64_t N = t->n;
  const int nq = s->nq;
  const double eps = p->epsilon;
  const double epsilon_inv = 1.0 / eps;
  const int use_lse = p->use_log_sum_exp;
  const int nthr = orc_threads(p);
  const int faithful = p->faithful_copier;

  double *tgrad = (double *)calloc((size_t)nthr * (size_t)N, sizeof(double));
  double *tF = (double *)calloc((size_t)nthr, sizeof(double));
  double *tC = (double *)calloc((size_t)nthr, sizeof(double));
  if (!tgrad || !tF || !tC) {
    free(tgrad); free(tF); free(tC);
    return -1;
  }

#pragma omp parallel num_threads(nthr)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
    double *acc = tgrad + (size_t)tid * N;
    double *copy_grad = faithful ? (double *)calloc((size_t)N, sizeof(double)) : acc;
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (N > 0 ? N : 1));
    double *ty = (double *)malloc(sizeof(double) * (N > 0 ? N : 1) * dim);
    double *tn = (double *)malloc(sizeof(double) * (N > 0 ? N : 1));
    double *tp = (double *)malloc(sizeof(double) * (N > 0 ? N : 1));
    double *exp_terms = (double *)malloc(sizeof(double) * (N > 0 ? N : 1));
    double myF = 0.0, myC = 0.0;

#pragma omp for schedule(static)
    for (int64_t c = 0; c < s->n_cells; ++c) {
      double cell_F = 0.0, cell_C = 0.0;
      if (faithful) memset(copy_grad, 0, sizeof(double) * (size_t)N);
      int64_t nt = orc_candidates(t, s->center + c * dim, dim, R, idx);
      if (nt > 0) {
        for (int64_t i = 0; i < nt; ++i) {
          int64_t j = idx[i];
          for (int d = 0; d < dim; ++d) ty[i * dim + d] = t->y[j * dim + d];
          tn[i] = t->nu[j];
          tp[i] = psi[j];
        }
        for (int q = 0; q < nq; ++q) {
          const int64_t g = c * nq + q;
          const double *xq = s->x + g * dim;
          const double density_value = s->rho[g];
          const double JxW = s->jxw[g];
          double total_sum_exp = 0.0;
          double max_exponent = -DBL_MAX;
          if (use_lse) {
            for (int64_t i = 0; i < nt; ++i) {
              double local_dist2 = pow(orc_dist(xq, ty + i * dim, dim), 2);
              double exponent = (tp[i] - 0.5 * local_dist2) * epsilon_inv;
              max_exponent = exponent > max_exponent ? exponent : max_exponent;
            }
            for (int64_t i = 0; i < nt; ++i) {
              double local_dist2 = pow(orc_dist(xq, ty + i * dim, dim), 2);
              double shifted_exp = exp((tp[i] - 0.5 * local_dist2) * epsilon_inv - max_exponent);
              exp_terms[i] = tn[i] * shifted_exp;
              total_sum_exp += exp_terms[i];
            }
          } else {
            for (int64_t i = 0; i < nt; ++i) {
              double local_dist2 = pow(orc_dist(xq, ty + i * dim, dim), 2);
              exp_terms[i] = tn[i] * exp((tp[i] - 0.5 * local_dist2) * epsilon_inv);
              total_sum_exp += exp_terms[i];
            }
          }
          if (total_sum_exp <= 0.0) continue;
          if (use_lse)
            cell_F += density_value * eps * (max_exponent + log(total_sum_exp)) * JxW;
          else
            cell_F += density_value * eps * log(total_sum_exp) * JxW;
          const double scale = density_value * JxW / total_sum_exp;
          cell_C += scale;
          for (int64_t i = 0; i < nt; ++i)
            if (exp_terms[i] > 0.0) copy_grad[idx[i]] += scale * exp_terms[i];
        }
      }
      /* copier (SotSolver.cc:307-311) */
      myF += cell_F;
      myC += cell_C;
      if (faithful)
        for (int64_t j = 0; j < N; ++j) acc[j] += copy_grad[j];
    }
    tF[tid] = myF;
    tC[tid] = myC;
    if (faithful) free(copy_grad);
    free(idx); free(ty); free(tn); free(tp); free(exp_terms);
  }

  double Ft = 0.0, Ct = 0.0;
  for (int k = 0; k < nthr; ++k) {
    Ft += tF[k];
    Ct += tC[k];
  }
  for (int64_t j = 0; j < N; ++j) {
    double gsum = 0.0;
    for (int k = 0; k < nthr; ++k) gsum += tgrad[(size_t)k * N + j];
    grad[j] = gsum;
  }
  /* linear term (SotSolver.cc:321-327) */
  for (int64_t j = 0; j < N; ++j) {
    Ft -= psi[j] * t->nu[j];
    grad[j] -= t->nu[j];
  }
  free(tgrad); free(tF); free(tC);
  *F = Ft;
  if (C) *C = Ct;
  return 0;
}

int orc_evaluate(const orc_source_t *s, const orc_target_t *t,
                 const orc_params_t *p, orc_state_t *st, const double *psi,
                 double *grad, double *F) {
  double R = orc_distance_threshold(p, st, s, t, psi);
  st->last_threshold = R;
  double C = 0.0;
  int rc = orc_evaluate_R(s, t, p, R, psi, grad, F, &C);
  st->global_functional = *F;
  st->C_global = C;
  return rc;
}

/* Discarded mass of the truncation (what the pointwise/integral/geometric
 * bounds of SotSolver.cc:379-551 are meant to keep below tau): for each listed
 * quadrature point, the fraction of its untruncated mass
 * sum_j nu_j exp(e_qj) that falls on targets outside its cell's candidate set. */
int orc_discarded_fraction(const orc_source_t *s, const orc_target_t *t,
                           const orc_params_t *p, double R, const double *psi,
                           int64_t n_list, const int64_t *q_list, double *frac) {
  const int dim = s->dim;
  const double inv = 1.0 / p->epsilon;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t k = 0; k < n_list; ++k) {
    const int64_t g = q_list[k];
    const int64_t c = g / s->nq;
    const double *xq = s->x + g * dim;
    double mx = -DBL_MAX;
    for (int64_t j = 0; j < t->n; ++j) {
      double d2 = pow(orc_dist(xq, t->y + j * dim, dim), 2);
      double e = (psi[j] - 0.5 * d2) * inv;
      mx = e > mx ? e : mx;
    }
    double all = 0.0, out = 0.0;
    for (int64_t j = 0; j < t->n; ++j) {
      double d2 = pow(orc_dist(xq, t->y + j * dim, dim), 2);
      double v = t->nu[j] * exp((psi[j] - 0.5 * d2) * inv - mx);
      all += v;
      if (!(orc_dist(t->y + j * dim, s->center + c * dim, dim) <= R)) out += v;
    }
    frac[k] = all > 0 ? out / all : 0.0;
  }
  return 0;
}

/* Dense Hessian of the (truncated) dual functional.  compute_hessian is only
 * declared in the reference (SotSolver.h:338-344); the matrix is the
 * derivative of the gradient assembled above:
 *   H_jk = (1/eps) sum_q w_q (delta_jk p_qj - p_qj p_qk),
 *   p_qj = exp_terms_j / total_sum_exp. */
int orc_hessian(const orc_source_t *s, const orc_target_t *t,
                const orc_params_t *p, double R, const double *psi, double *H) {
  const int dim = s->dim;
  const int64_t N = t->n;
  const int nq = s->nq;
  const double eps = p->epsilon, epsilon_inv = 1.0 / eps;
  const int nthr = orc_threads(p);
  double *tH = (double *)calloc((size_t)nthr * N * N, sizeof(double));
  if (!tH) return -1;
#pragma omp parallel num_threads(nthr)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
    double *Hl = tH + (size_t)tid * N * N;
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (N > 0 ? N : 1));
    double *pe = (double *)malloc(sizeof(double) * (N > 0 ? N : 1));
#pragma omp for schedule(static)
    for (int64_t c = 0; c < s->n_cells; ++c) {
      int64_t nt = orc_candidates(t, s->center + c * dim, dim, R, idx);
      for (int q = 0; q < nq && nt > 0; ++q) {
        const int64_t g = c * nq + q;
        const double *xq = s->x + g * dim;
        double mx = p->use_log_sum_exp ? -DBL_MAX : 0.0;
        if (p->use_log_sum_exp)
          for (int64_t i = 0; i < nt; ++i) {
            double d2 = pow(orc_dist(xq, t->y + idx[i] * dim, dim), 2);
            double e = (psi[idx[i]] - 0.5 * d2) * epsilon_inv;
            mx = e > mx ? e : mx;
          }
        double S = 0.0;
        for (int64_t i = 0; i < nt; ++i) {
          double d2 = pow(orc_dist(xq, t->y + idx[i] * dim, dim), 2);
          pe[i] = t->nu[idx[i]] * exp((psi[idx[i]] - 0.5 * d2) * epsilon_inv - mx);
          S += pe[i];
        }
        if (S <= 0.0) continue;
        const double w = s->rho[g] * s->jxw[g];
        for (int64_t i = 0; i < nt; ++i) pe[i] /= S;
        for (int64_t a = 0; a < nt; ++a) {
          const double wa = w * pe[a] * epsilon_inv;
          double *row = Hl + (size_t)idx[a] * N;
          row[idx[a]] += wa;
          for (int64_t b = 0; b < nt; ++b) row[idx[b]] -= wa * pe[b];
        }
      }
    }
    free(idx);
    free(pe);
  }
  for (int64_t k = 0; k < N * N; ++k) {
    double v = 0.0;
    for (int th = 0; th < nthr; ++th) v += tH[(size_t)th * N * N + k];
    H[k] = v;
  }
  free(tH);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* compute_weighted_barycenters_euclidean (SotSolver.cc:871-971): the          */
/* local_assemble loop with the barycenter function_call (:912-934), copier    */
/* sum, then division by nu_j (:957-963).                                     */
/* ------------------------------------------------------------------------ */

int orc_barycenters_R(const orc_source_t *s, const orc_target_t *t,
                      const orc_params_t *p, double R, const double *psi,
                      double *bary) {
  const int dim = s->dim;
  const int64_t N = t->n;
  const int nq = s->nq;
  const double epsilon_inv = 1.0 / p->epsilon;
  const int use_lse = p->use_log_sum_exp;
  const int nthr = orc_threads(p);
  double *tb = (double *)calloc((size_t)nthr * (size_t)N * dim, sizeof(double));
  if (!tb) return -1;
#pragma omp parallel num_threads(nthr)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
    double *acc = tb + (size_t)tid * N * dim;
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (N > 0 ? N : 1));
    double *exp_terms = (double *)malloc(sizeof(double) * (N > 0 ? N : 1));
#pragma omp for schedule(static)
    for (int64_t c = 0; c < s->n_cells; ++c) {
      const int64_t nt = orc_candidates(t, s->center + c * dim, dim, R, idx);
      if (nt == 0) continue;
      for (int q = 0; q < nq; ++q) {
        const int64_t g = c * nq + q;
        const double *xq = s->x + g * dim;
        double total_sum_exp = 0.0, max_exponent = -DBL_MAX;
        if (use_lse)
          for (int64_t i = 0; i < nt; ++i) {
            const int64_t j = idx[i];
            const double d2 = pow(orc_dist(xq, t->y + j * dim, dim), 2);
            const double e = (psi[j] - 0.5 * d2) * epsilon_inv;
            max_exponent = e > max_exponent ? e : max_exponent;
          }
        for (int64_t i = 0; i < nt; ++i) {
          const int64_t j = idx[i];
          const double d2 = pow(orc_dist(xq, t->y + j * dim, dim), 2);
          const double e = (psi[j] - 0.5 * d2) * epsilon_inv;
          exp_terms[i] = t->nu[j] * (use_lse ? exp(e - max_exponent) : exp(e));
          total_sum_exp += exp_terms[i];
        }
        if (total_sum_exp <= 0.0) continue;
        const double scale = s->rho[g] * s->jxw[g] / total_sum_exp;
        for (int64_t i = 0; i < nt; ++i)
          if (exp_terms[i] > 0.0)
            for (int d = 0; d < dim; ++d) acc[dim * idx[i] + d] += scale * exp_terms[i] * xq[d];
      }
    }
    free(idx);
    free(exp_terms);
  }
  for (int64_t j = 0; j < N; ++j)
    for (int d = 0; d < dim; ++d) {
      double v = 0.0;
      for (int k = 0; k < nthr; ++k) v += tb[(size_t)k * N * dim + dim * j + d];
      bary[j * dim + d] = v / t->nu[j];
    }
  free(tb);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* get_potential_conditioned_density (SotSolver.cc:973-1046): at every point  */
/* (a DoF support point), the LSE over the targets within R of it; for each   */
/* requested target i among them, rho_p exp(e_pi - M_p) / sum_j nu_j           */
/* exp(e_pj - M_p).  out [n_idx * n_points], 0 where i is not a candidate.     */
/* ------------------------------------------------------------------------ */

int orc_conditional_density(const orc_target_t *t, int dim, const double *psi,
                            double epsilon, double R, int64_t n_points,
                            const double *points, const double *density,
                            int64_t n_idx, const int64_t *indices, double *out) {
  const int64_t N = t->n;
  const double epsilon_inv = 1.0 / epsilon;
  memset(out, 0, sizeof(double) * (size_t)(n_idx * n_points));
#pragma omp parallel
  {
    int64_t *cand = (int64_t *)malloc(sizeof(int64_t) * (N > 0 ? N : 1));
    char *is_cand = (char *)calloc((size_t)(N > 0 ? N : 1), 1);
    double *ex = (double *)calloc((size_t)(N > 0 ? N : 1), sizeof(double));
#pragma omp for schedule(static)
    for (int64_t p = 0; p < n_points; ++p) {
      const double *sp = points + p * dim;
      const int64_t nt = orc_candidates(t, sp, dim, R, cand);
      double total_sum_exp = 0.0, max_exponent = -DBL_MAX;
      for (int64_t i = 0; i < nt; ++i) {
        const int64_t j = cand[i];
        const double e = (psi[j] - 0.5 * pow(orc_dist(sp, t->y + j * dim, dim), 2)) * epsilon_inv;
        max_exponent = e > max_exponent ? e : max_exponent;
      }
      for (int64_t i = 0; i < nt; ++i) {
        const int64_t j = cand[i];
        ex[j] = exp((psi[j] - 0.5 * pow(orc_dist(sp, t->y + j * dim, dim), 2)) * epsilon_inv -
                    max_exponent);
        total_sum_exp += t->nu[j] * ex[j];
        is_cand[j] = 1;
      }
      if (total_sum_exp > 0.0)
        for (int64_t k = 0; k < n_idx; ++k) {
          const int64_t i = indices[k];
          if (i >= 0 && i < N && is_cand[i]) out[k * n_points + p] = density[p] * (ex[i] / total_sum_exp);
        }
      for (int64_t i = 0; i < nt; ++i) {
        is_cand[cand[i]] = 0;
        ex[cand[i]] = 0.0;
      }
    }
    free(cand);
    free(is_cand);
    free(ex);
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* ModalStrategy / BarycentricStrategy::compute_map (OptimalTransportPlan.cc: */
/* 99-189, 218-310).  Candidates of x: the targets with dist(x, y) < radius   */
/* when radius > 0 (all otherwise), in index order; none: the nearest target  */
/* (lowest index among equals).                                               */
/*   modal:       the target maximising psi_j - |x-y_j|^2/2 + eps log nu_j    */
/*                (first maximum in candidate order);                         */
/*   barycentric: sum_j nu_j e^(e_j - M) y_j / sum_j nu_j e^(e_j - M),         */
/*                e_j = (psi_j - |x-y_j|^2/2) / eps; the nearest target when  */
/*                the sum is not positive.                                    */
/* ------------------------------------------------------------------------ */

static int64_t orc_nearest(const orc_target_t *t, const double *x, int dim) {
  int64_t best = 0;
  double bd = INFINITY;
  for (int64_t j = 0; j < t->n; ++j) {
    const double d = orc_dist(x, t->y + j * dim, dim);
    if (d < bd) {
      bd = d;
      best = j;
    }
  }
  return best;
}

int orc_transport_map(int strategy, const orc_target_t *t, int dim,
                      const double *psi, double epsilon, double radius,
                      int64_t n_points, const double *points, double *mapped) {
  const int64_t N = t->n;
  const int use_truncation = radius > 0.0;
#pragma omp parallel
  {
    int64_t *cand = (int64_t *)malloc(sizeof(int64_t) * (N > 0 ? N : 1));
    double *et = (double *)malloc(sizeof(double) * (N > 0 ? N : 1));
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n_points; ++i) {
      const double *x = points + i * dim;
      int64_t nc = 0;
      if (use_truncation) {
        for (int64_t j = 0; j < N; ++j)
          if (orc_dist(x, t->y + j * dim, dim) < radius) cand[nc++] = j;
        if (nc == 0) {
          const int64_t j = orc_nearest(t, x, dim);
          for (int d = 0; d < dim; ++d) mapped[i * dim + d] = t->y[j * dim + d];
          continue;
        }
      } else {
        for (int64_t j = 0; j < N; ++j) cand[nc++] = j;
      }
      if (strategy == 0) {
        double max_score = -INFINITY;
        int64_t best = 0;
        for (int64_t k = 0; k < nc; ++k) {
          const int64_t j = cand[k];
          const double sq = pow(orc_dist(x, t->y + j * dim, dim), 2);
          const double log_term = t->nu[j] > 0 ? epsilon * log(t->nu[j]) : -INFINITY;
          const double score = psi[j] - 0.5 * sq + log_term;
          if (score > max_score) {
            max_score = score;
            best = j;
          }
        }
        for (int d = 0; d < dim; ++d) mapped[i * dim + d] = t->y[best * dim + d];
      } else {
        double max_exponent = -INFINITY;
        for (int64_t k = 0; k < nc; ++k) {
          const int64_t j = cand[k];
          const double sq = pow(orc_dist(x, t->y + j * dim, dim), 2);
          et[k] = (psi[j] - 0.5 * sq) / epsilon;
          max_exponent = et[k] > max_exponent ? et[k] : max_exponent;
        }
        double ws[3] = {0.0, 0.0, 0.0}, sum_exp = 0.0;
        for (int64_t k = 0; k < nc; ++k) {
          const int64_t j = cand[k];
          const double weight = t->nu[j] * exp(et[k] - max_exponent);
          for (int d = 0; d < dim; ++d) ws[d] += weight * t->y[j * dim + d];
          sum_exp += weight;
        }
        if (sum_exp > 0) {
          for (int d = 0; d < dim; ++d) mapped[i * dim + d] = ws[d] / sum_exp;
        } else {
          const int64_t j = orc_nearest(t, x, dim);
          for (int d = 0; d < dim; ++d) mapped[i * dim + d] = t->y[j * dim + d];
        }
      }
    }
    free(cand);
    free(et);
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* SoftmaxRefinement::compute_refinement (SoftmaxRefinement.cc:200-280) and    */
/* its local_assemble (SoftmaxRefinement.cc:61-198): the potential of a fine   */
/* target level from the coarse level's potential,                           */
/*   psi_f = -lambda log sum_q (w_q / Z_q) e^{-M_q} exp(-|x_q - y_f|^2/(2 lambda)),*/
/*   Z_q e^{M_q} = sum_c nu_c exp((psi_c - |x_q - y_c|^2/2)/lambda),             */
/* where c runs over the coarse candidates of the cell of q (cell rule by the  */
/* cell center, :81, :43-59) that pass the per-q test d^2 <= R^2 (:152,:162),  */
/* and f over the children of those candidates that pass the same test (:99-  */
/* 128, :192).  Fine points that receive nothing keep psi_f = 0 (:270).       */
/* ------------------------------------------------------------------------ */

int orc_softmax_refinement(const orc_source_t *s, int64_t n_fine,
                           const double *y_fine, const orc_target_t *coarse,
                           const double *psi_coarse, int64_t n_parents,
                           const int64_t *child_off, const int64_t *child_idx,
                           double lambda, double threshold, int use_lse,
                           int n_threads, double *psi_fine) {
  const int dim = s->dim;
  const int64_t Nc = coarse->n;
  const int nq = s->nq;
  const double lambda_inv = 1.0 / lambda;
  const double threshold_sq = threshold * threshold;
  orc_params_t tp;
  memset(&tp, 0, sizeof(tp));
  tp.n_threads = n_threads;
  const int nthr = orc_threads(&tp);
  const int64_t nf_alloc = n_fine > 0 ? n_fine : 1;
  double *tpot = (double *)calloc((size_t)nthr * (size_t)nf_alloc, sizeof(double));
  if (!tpot) return -1;

#pragma omp parallel num_threads(nthr)
  {
#ifdef _OPENMP
    const int tid = omp_get_thread_num();
#else
    const int tid = 0;
#endif
    double *pot = tpot + (size_t)tid * nf_alloc;
    int64_t *cidx = (int64_t *)malloc(sizeof(int64_t) * (Nc > 0 ? Nc : 1));
    int64_t cap = 64, nfl = 0;
    int64_t *fidx = (int64_t *)malloc(sizeof(int64_t) * cap);
#pragma omp for schedule(static)
    for (int64_t c = 0; c < s->n_cells; ++c) {
      /* coarse candidates of the cell (:81) */
      int64_t nt = orc_candidates(coarse, s->center + c * dim, dim, threshold, cidx);
      if (nt == 0) continue;
      /* children of the candidates, in candidate order (:109-128) */
      nfl = 0;
      for (int64_t i = 0; i < nt; ++i) {
        const int64_t ci = cidx[i];
        if (ci >= n_parents) continue;
        for (int64_t k = child_off[ci]; k < child_off[ci + 1]; ++k) {
          const int64_t f = child_idx[k];
          if (f < 0 || f >= n_fine) continue;
          if (nfl == cap) {
            cap *= 2;
            fidx = (int64_t *)realloc(fidx, sizeof(int64_t) * cap);
          }
          fidx[nfl++] = f;
        }
      }
      if (nfl == 0) continue;
      for (int q = 0; q < nq; ++q) {
        const int64_t g = c * nq + q;
        const double *x = s->x + g * dim;
        const double density_value = s->rho[g];
        const double JxW = s->jxw[g];
        double total_sum_exp = 0.0;
        double max_exponent = -DBL_MAX;
        if (use_lse) {
          for (int64_t i = 0; i < nt; ++i) {
            const int64_t j = cidx[i];
            const double local_dist2 = pow(orc_dist(x, coarse->y + j * dim, dim), 2);
            if (local_dist2 <= threshold_sq) {
              const double exponent = (psi_coarse[j] - 0.5 * local_dist2) * lambda_inv;
              max_exponent = exponent > max_exponent ? exponent : max_exponent;
            }
          }
          for (int64_t i = 0; i < nt; ++i) {
            const int64_t j = cidx[i];
            const double local_dist2 = pow(orc_dist(x, coarse->y + j * dim, dim), 2);
            if (local_dist2 <= threshold_sq) {
              const double shifted_exp =
                  exp((psi_coarse[j] - 0.5 * local_dist2) * lambda_inv - max_exponent);
              total_sum_exp += coarse->nu[j] * shifted_exp;
            }
          }
        } else {
          for (int64_t i = 0; i < nt; ++i) {
            const int64_t j = cidx[i];
            const double local_dist2 = pow(orc_dist(x, coarse->y + j * dim, dim), 2);
            if (local_dist2 <= threshold_sq)
              total_sum_exp += coarse->nu[j] * exp((psi_coarse[j] - 0.5 * local_dist2) * lambda_inv);
          }
        }
        if (total_sum_exp <= 0.0) continue;
        double scale = density_value * JxW / total_sum_exp;
        if (use_lse) scale *= exp(-max_exponent);
        for (int64_t i = 0; i < nfl; ++i) {
          const int64_t f = fidx[i];
          const double local_dist2_fine = pow(orc_dist(x, y_fine + f * dim, dim), 2);
          if (local_dist2_fine <= threshold_sq)
            pot[f] += scale * exp((-0.5 * local_dist2_fine) * lambda_inv);
        }
      }
    }
    free(cidx);
    free(fidx);
  }
  /* copier sum, then psi_f = -lambda log(pot_f) where pot_f > 0 (:264-274) */
  for (int64_t f = 0; f < n_fine; ++f) {
    double v = 0.0;
    for (int k = 0; k < nthr; ++k) v += tpot[(size_t)k * nf_alloc + f];
    psi_fine[f] = v > 0.0 ? -lambda * log(v) : v;
  }
  free(tpot);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* L-BFGS (deal.II SolverBFGS, m = 5, no preconditioner) and its default line */
/* search (deal.II LineMinimization::line_search with poly_fit, eta=0.9,      */
/* mu=0.001, at most 20 evaluations).  The in-tree text of both algorithms is */
/* core/wasserstein_solver_bfgs.h:300-492 and core/line_minimization.h:       */
/* 332-455, 498-749; SotSolver::solve wires them at SotSolver.cc:105-215.     */
/* ------------------------------------------------------------------------ */

typedef double (*orc_eval_fn)(void *ctx, const double *x, double *g);

static double dotv(const double *a, const double *b, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

static int cubic_fit(double x1, double f1, double g1, double x2, double f2,
                     double g2, double *out) {
  const double beta1 = g1 + g2 - 3. * (f1 - f2) / (x1 - x2);
  const double s = beta1 * beta1 - g1 * g2;
  if (s < 0) return 0;
  const double beta2 = sqrt(s);
  const double denom = x1 < x2 ? g2 - g1 + 2. * beta2 : g1 - g2 + 2. * beta2;
  if (denom == 0.) return 0;
  *out = x1 < x2 ? x2 - (x2 - x1) * (g2 + beta2 - beta1) / denom
                 : x1 - (x1 - x2) * (g1 + beta2 - beta1) / denom;
  return 1;
}

static int quadratic_fit(double x1, double f1, double g1, double x2, double f2,
                         double *out) {
  const double denom = (2. * g1 * x2 - 2. * g1 * x1 - 2. * f2 + 2. * f1);
  if (denom == 0) return 0;
  *out = (g1 * (x2 * x2 - x1 * x1) + 2. * (f1 - f2) * x1) / denom;
  return 1;
}

static double poly_fit(double x1, double f1, double g1, double x2, double f2,
                       double g2, double blo, double bhi) {
  double r;
  if (cubic_fit(x1, f1, g1, x2, f2, g2, &r) && r >= blo && r <= bhi) return r;
  if (quadratic_fit(x1, f1, g1, x2, f2, &r) && r >= blo && r <= bhi) return r;
  return (blo + bhi) * 0.5;
}

typedef struct {
  orc_eval_fn fn;
  void *ctx;
  int64_t n;
  double *x, *x0, *g;
  const double *p;
  double f;
  int evals;
} ls_ctx;

static void ls_func(ls_ctx *c, double a, double *fo, double *go) {
  for (int64_t i = 0; i < c->n; ++i) c->x[i] = c->x0[i] + a * c->p[i];
  c->f = c->fn(c->ctx, c->x, c->g);
  c->evals++;
  *fo = c->f;
  *go = dotv(c->g, c->p, c->n);
}

/* returns 0 and the step, or -1 when the bracket or sectioning phase runs out
 * of evaluations (stock deal.II throws there). */
static int line_search(ls_ctx *c, double f0, double g0, double a1, double eta,
                       double mu, double a_max, unsigned max_evaluations,
                       double *alpha) {
  const double tau1 = 9., tau2 = 0.1, tau3 = 0.5;
  const double g0_abs = fabs(g0);
  const double f_min = f0 + a_max * mu * g0;
#define W1(a, f) ((f) <= f0 + (a)*mu * g0)
#define W2(g) (fabs(g) <= eta * g0_abs)
  double a_lo = NAN, f_lo = NAN, g_lo = NAN, a_hi = NAN, f_hi = NAN, g_hi = NAN;
  double ai, fi, gi;
  unsigned i = 0;
  {
    double f_prev = f0, g_prev = g0, a_prev = 0;
    ai = a1;
    while (i < max_evaluations) {
      ls_func(c, ai, &fi, &gi);
      ++i;
      if (fi <= f_min || ai == a_max) {
        *alpha = ai;
        return 0;
      }
      if (!W1(ai, fi) || (fi >= f_prev && i > 1)) {
        a_lo = a_prev; f_lo = f_prev; g_lo = g_prev;
        a_hi = ai; f_hi = fi; g_hi = gi;
        break;
      }
      if (W2(gi)) {
        *alpha = ai;
        return 0;
      }
      if (gi >= 0) {
        a_lo = ai; f_lo = fi; g_lo = gi;
        a_hi = a_prev; f_hi = f_prev; g_hi = g_prev;
        break;
      }
      double up = ai + tau1 * (ai - a_prev);
      a_prev = ai; f_prev = fi; g_prev = gi;
      ai = a_max < up ? a_max : up;
    }
  }
  if (!(i < max_evaluations)) return -1;
  while (i < max_evaluations) {
    const double a_lo_safe = a_lo + tau2 * (a_hi - a_lo);
    const double a_hi_safe = a_hi - tau3 * (a_hi - a_lo);
    const double blo = a_lo_safe < a_hi_safe ? a_lo_safe : a_hi_safe;
    const double bhi = a_lo_safe < a_hi_safe ? a_hi_safe : a_lo_safe;
    ai = poly_fit(a_lo, f_lo, g_lo, a_hi, f_hi, g_hi, blo, bhi);
    ls_func(c, ai, &fi, &gi);
    ++i;
    if (!W1(ai, fi) || fi >= f_lo) {
      a_hi = ai; f_hi = fi; g_hi = gi;
    } else {
      if (W2(gi)) {
        *alpha = ai;
        return 0;
      }
      if (gi * (a_hi - a_lo) >= 0) {
        a_hi = a_lo; f_hi = f_lo; g_hi = g_lo;
      }
      a_lo = ai; f_lo = fi; g_lo = gi;
    }
  }
#undef W1
#undef W2
  return -1;
}

typedef int (*orc_check_fn)(void *ctx, int step, const double *g, double *check_value);

/* SolverBFGS::solve; check() plays SolverControl::check (success when
 * check_value <= tol, failure at step >= max or NaN). */
static int bfgs_core(orc_eval_fn fn, void *ctx, orc_check_fn chk, void *cctx,
                     int64_t n, double *x, orc_solve_info_t *info) {
  const int m_max = 5;
  double *g = (double *)calloc(n, sizeof(double));
  double *p = (double *)calloc(n, sizeof(double));
  double *yk = (double *)calloc(n, sizeof(double));
  double *sk = (double *)calloc(n, sizeof(double));
  double *x0 = (double *)calloc(n, sizeof(double));
  double *S = (double *)calloc((size_t)m_max * n, sizeof(double));
  double *Y = (double *)calloc((size_t)m_max * n, sizeof(double));
  double rho[5], c1[5];
  int order[5]; /* order[0] = newest slot */
  int m = 0;
  int status;
  ls_ctx lc = {fn, ctx, n, x, x0, g, p, 0.0, 0};
  double f = fn(ctx, x, g);
  lc.evals = 1;
  int k = 0;
  double cv;
  int st = chk(cctx, k, g, &cv);
  int first_step = 1;
  double f_prev = 0.;
  while (st == 0) {
    for (int64_t i = 0; i < n; ++i) p[i] = g[i];
    for (int i = 0; i < m; ++i) {
      const double *si = S + (size_t)order[i] * n, *yi = Y + (size_t)order[i] * n;
      c1[i] = rho[order[i]] * dotv(si, p, n);
      for (int64_t t = 0; t < n; ++t) p[t] += -c1[i] * yi[t];
    }
    for (int i = m - 1; i >= 0; --i) {
      const double *si = S + (size_t)order[i] * n, *yi = Y + (size_t)order[i] * n;
      const double c2 = rho[order[i]] * dotv(yi, p, n);
      for (int64_t t = 0; t < n; ++t) p[t] += (c1[i] - c2) * si[t];
    }
    for (int64_t i = 0; i < n; ++i) p[i] *= -1.;
    memcpy(sk, x, sizeof(double) * n);
    memcpy(yk, g, sizeof(double) * n);
    /* default line minimisation */
    {
      const double f0 = f;
      const double g0 = dotv(g, p, n);
      memcpy(x0, x, sizeof(double) * n);
      double df = f_prev - f;
      df = df > 100. * DBL_EPSILON ? df : 100. * DBL_EPSILON;
      double a1;
      if (first_step) {
        a1 = 1.;
      } else {
        double t = -1.01 * 2. * df / g0;
        a1 = 1. < t ? 1. : t;
      }
      f_prev = f;
      double alpha;
      lc.f = f;
      int rc = line_search(&lc, f0, g0, a1, 0.9, 0.001, DBL_MAX, 20, &alpha);
      f = lc.f;
      if (rc != 0) {
        status = ORC_LINESEARCH_FAIL;
        k++;
        goto done;
      }
      first_step = 0;
    }
    for (int64_t i = 0; i < n; ++i) {
      sk[i] = -sk[i] + x[i];
      yk[i] = -yk[i] + g[i];
    }
    ++k;
    st = chk(cctx, k, g, &cv);
    if (st != 0) break;
    {
      const double curvature = dotv(sk, yk, n);
      if (curvature > 0.) {
        int slot;
        if (m < m_max) {
          slot = m;
          for (int i = m; i > 0; --i) order[i] = order[i - 1];
          m++;
        } else {
          slot = order[m_max - 1];
          for (int i = m_max - 1; i > 0; --i) order[i] = order[i - 1];
        }
        order[0] = slot;
        memcpy(S + (size_t)slot * n, sk, sizeof(double) * n);
        memcpy(Y + (size_t)slot * n, yk, sizeof(double) * n);
        rho[slot] = 1. / curvature;
      }
    }
  }
  status = st == 1 ? ORC_CONVERGED : (isnan(cv) ? ORC_NONFINITE : ORC_MAX_ITER);
done:
  if (info) {
    info->status = status;
    info->iterations = k;
    info->evaluations = lc.evals;
    info->functional = f;
    info->check_value = cv;
  }
  free(g); free(p); free(yk); free(sk); free(x0); free(S); free(Y);
  return status;
}

/* ---- SOT wiring (SotSolver::solve, SotSolver.cc:105-215) ---- */
typedef struct {
  const orc_source_t *s;
  const orc_target_t *t;
  const orc_params_t *p;
  orc_state_t *st;
} sot_ctx;

static double sot_eval(void *vc, const double *x, double *g) {
  sot_ctx *c = (sot_ctx *)vc;
  double F = 0.0;
  orc_evaluate(c->s, c->t, c->p, c->st, x, g, &F);
  return F;
}

/* VerboseSolverControl::check (SotSolver.h:382-446) + SolverControl::check */
typedef struct {
  const orc_params_t *p;
  const orc_target_t *t;
} sot_chk_ctx;

static int sot_check(void *vc, int step, const double *g, double *cvo) {
  sot_chk_ctx *c = (sot_chk_ctx *)vc;
  const int64_t n = c->t->n;
  double cv, tol;
  if (c->p->control_type == ORC_CTRL_COMPONENTWISE) {
    cv = -INFINITY;
    for (int64_t j = 0; j < n; ++j) {
      double r = fabs(g[j]) - c->t->nu[j] * c->p->tolerance;
      cv = r > cv ? r : cv;
    }
    tol = DBL_MIN;
  } else {
    cv = 0.0;
    for (int64_t j = 0; j < n; ++j) cv += fabs(g[j]);
    tol = c->p->tolerance;
  }
  *cvo = cv;
  if (cv <= tol) return 1;
  if (step >= c->p->max_iterations || isnan(cv)) return 2;
  return 0;
}

int orc_solve(const orc_source_t *s, const orc_target_t *t,
              const orc_params_t *p, orc_state_t *st, double *psi,
              orc_solve_info_t *info) {
  if (t->n <= 0) return -1;
  st->covering_radius = orc_covering_radius(s, t) * 1.1;
  st->min_target_density = vec_min(t->nu, t->n);
  sot_ctx c = {s, t, p, st};
  sot_chk_ctx cc = {p, t};
  return bfgs_core(sot_eval, &c, sot_check, &cc, t->n, psi, info);
}

/* ---- generic objective (pins the BFGS text on textbook problems) ---- */
typedef struct {
  double tol;
  int max_it;
  int64_t n;
} gen_chk_ctx;

static int gen_check(void *vc, int step, const double *g, double *cvo) {
  gen_chk_ctx *c = (gen_chk_ctx *)vc;
  double cv = sqrt(dotv(g, g, c->n));
  *cvo = cv;
  if (cv <= c->tol) return 1;
  if (step >= c->max_it || isnan(cv)) return 2;
  return 0;
}

int orc_bfgs_generic(orc_fun_t fun, void *ctx, int64_t n, double *x,
                     int max_iterations, double tol_l2, orc_solve_info_t *info) {
  gen_chk_ctx cc = {tol_l2, max_iterations, n};
  return bfgs_core((orc_eval_fn)fun, ctx, gen_check, &cc, n, x, info);
}
