// sdot_mesh.cpp -- source-measure extraction: the flat arrays deal.II's
// FEValues would hand to SotSolver::local_assemble (SotSolver.cc:788-790):
// quadrature points, rho(x_q) * JxW_q and cell centers, for
//   * structured Q1 boxes (GridGenerator::subdivided_hyper_rectangle, FE_Q(1),
//     MappingQ1, QGauss(order)), and
//   * P1 simplices (FE_SimplexP(1), MappingFE, QGaussSimplex(order)),
//   * general Q1 quadrilaterals and hexahedra (FE_Q(1), MappingQ1,
//     QGauss(order)) with vertices in deal.II's lexicographic order, also as
//     surfaces in 3D (SotSolver<2, 3>), and P1 triangles in 3D,
// with the element/quadrature choice of utils.h:776-840 and the density
// normalisation of SemiDiscreteOT.cc:422-449.  For a cell of lower dimension
// than the space, JxW is the Gram measure sqrt(det(J^T J)) times the weight,
// as deal.II's FEValues computes it for codimension one.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/sdot.h"

namespace {

// Gauss-Legendre rule on [0,1] by Newton's method on P_n.
void gauss01(int n, double *pts, double *wts) {
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
    double deriv = 0.0;
    for (int it = 0; it < 100; ++it) {
      double pn = 1.0, pm = 0.0;
      for (int k = 1; k <= n; ++k) {
        const double pk = ((2.0 * k - 1.0) * z * pn - (k - 1.0) * pm) / k;
        pm = pn;
        pn = pk;
      }
      deriv = n * (z * pn - pm) / (z * z - 1.0);
      const double step = pn / deriv;
      z -= step;
      if (std::fabs(step) < 1e-17) break;
    }
    double pn = 1.0, pm = 0.0;
    for (int k = 1; k <= n; ++k) {
      const double pk = ((2.0 * k - 1.0) * z * pn - (k - 1.0) * pm) / k;
      pm = pn;
      pn = pk;
    }
    deriv = n * (z * pn - pm) / (z * z - 1.0);
    pts[i] = 0.5 * (1.0 - z);  // z descends with i, so pts ascend
    wts[i] = 1.0 / ((1.0 - z * z) * deriv * deriv);
  }
}

struct Rule {
  int nq = 0;
  std::vector<double> xi;  // [nq*dim] reference coordinates
  std::vector<double> w;   // [nq]
};

Rule tensor_rule(int dim, int order) {
  double p[16], w[16];
  gauss01(order, p, w);
  Rule r;
  r.nq = 1;
  for (int d = 0; d < dim; ++d) r.nq *= order;
  r.xi.resize((size_t)r.nq * dim);
  r.w.resize(r.nq);
  for (int q = 0; q < r.nq; ++q) {
    int rem = q;
    double wq = 1.0;
    for (int d = 0; d < dim; ++d) {
      const int i = rem % order;
      rem /= order;
      r.xi[(size_t)q * dim + d] = p[i];
      wq *= w[i];
    }
    r.w[q] = wq;
  }
  return r;
}

bool simplex_rule(int dim, int order, Rule &r) {
  if (order == 1) {
    r.nq = 1;
    r.xi.assign(dim, 1.0 / (dim + 1));
    r.w.assign(1, dim == 2 ? 0.5 : 1.0 / 6.0);
    return true;
  }
  if (order == 2 && dim == 2) {
    const double a = 1.0 / 6.0, b = 2.0 / 3.0;
    r.nq = 3;
    r.xi = {a, a, b, a, a, b};
    r.w.assign(3, 1.0 / 6.0);
    return true;
  }
  if (order == 2 && dim == 3) {
    const double a = (5.0 - std::sqrt(5.0)) / 20.0, b = (5.0 + 3.0 * std::sqrt(5.0)) / 20.0;
    r.nq = 4;
    r.xi = {a, a, a, b, a, a, a, b, a, a, a, b};
    r.w.assign(4, 1.0 / 24.0);
    return true;
  }
  return false;
}

// Fills x, rho_q, jxw, centers of a structured box mesh.
void structured_fill(int dim, const int *ns, const double *lo, const double *hi, const Rule &r,
                     const double *nodal, double *x, double *rho, double *jxw, double *cen) {
  double h[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) h[d] = (hi[d] - lo[d]) / ns[d];
  double det = 1.0;
  for (int d = 0; d < dim; ++d) det *= h[d];
  const int nv = 1 << dim;
  const int64_t nx = ns[0], ny = ns[1], nz = dim == 3 ? ns[2] : 1;
  const int64_t vx = nx + 1, vy = ny + 1;
  const int64_t n_cells = nx * ny * nz;
  for (int64_t c = 0; c < n_cells; ++c) {
    const int64_t idx[3] = {c % nx, (c / nx) % ny, c / (nx * ny)};
    double V[8][3] = {};
    double Vr[8];
    for (int v = 0; v < nv; ++v) {
      int64_t vi[3] = {0, 0, 0};
      for (int d = 0; d < dim; ++d) {
        vi[d] = idx[d] + ((v >> d) & 1);
        V[v][d] = lo[d] + vi[d] * h[d];
      }
      const int64_t vid = vi[0] + vx * (vi[1] + vy * vi[2]);
      Vr[v] = nodal ? nodal[vid] : 1.0;
    }
    double cc[3] = {0, 0, 0};
    for (int v = 0; v < nv; ++v)
      for (int d = 0; d < dim; ++d) cc[d] += V[v][d];
    for (int d = 0; d < dim; ++d) cen[c * dim + d] = cc[d] / nv;
    for (int q = 0; q < r.nq; ++q) {
      const double *xi = &r.xi[(size_t)q * dim];
      double xp[3] = {0, 0, 0}, rv = 0.0;
      for (int v = 0; v < nv; ++v) {
        double phi = 1.0;
        for (int d = 0; d < dim; ++d) phi *= ((v >> d) & 1) ? xi[d] : 1.0 - xi[d];
        for (int d = 0; d < dim; ++d) xp[d] += V[v][d] * phi;
        rv += Vr[v] * phi;
      }
      const int64_t g = c * r.nq + q;
      for (int d = 0; d < dim; ++d) x[g * dim + d] = xp[d];
      rho[g] = rv;
      jxw[g] = det * r.w[q];
    }
  }
}

void simplex_fill(int dim, const double *vert, int64_t n_cells, const int64_t *cells,
                  const Rule &r, const double *nodal, double *x, double *rho, double *jxw,
                  double *cen) {
  const int nv = dim + 1;
  for (int64_t c = 0; c < n_cells; ++c) {
    double V[4][3] = {};
    double Vr[4];
    for (int i = 0; i < nv; ++i) {
      const int64_t vid = cells[c * nv + i];
      for (int d = 0; d < dim; ++d) V[i][d] = vert[vid * dim + d];
      Vr[i] = nodal ? nodal[vid] : 1.0;
    }
    double cc[3] = {0, 0, 0};
    for (int i = 0; i < nv; ++i)
      for (int d = 0; d < dim; ++d) cc[d] += V[i][d];
    for (int d = 0; d < dim; ++d) cen[c * dim + d] = cc[d] / nv;
    double J[3][3] = {};
    for (int d = 0; d < dim; ++d)
      for (int k = 0; k < dim; ++k) J[d][k] = V[k + 1][d] - V[0][d];
    double det = dim == 2 ? J[0][0] * J[1][1] - J[0][1] * J[1][0]
                          : J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    det = std::fabs(det);
    for (int q = 0; q < r.nq; ++q) {
      const double *xi = &r.xi[(size_t)q * dim];
      double phi[4];
      phi[0] = 1.0;
      for (int k = 0; k < dim; ++k) {
        phi[k + 1] = xi[k];
        phi[0] -= xi[k];
      }
      double xp[3] = {0, 0, 0}, rv = 0.0;
      for (int i = 0; i < nv; ++i) {
        for (int d = 0; d < dim; ++d) xp[d] += V[i][d] * phi[i];
        rv += Vr[i] * phi[i];
      }
      const int64_t g = c * r.nq + q;
      for (int d = 0; d < dim; ++d) x[g * dim + d] = xp[d];
      rho[g] = rv;
      jxw[g] = det * r.w[q];
    }
  }
}

// |det J| for a square Jacobian, sqrt(det(J^T J)) otherwise (J is
// spacedim x cdim, columns = derivatives along the reference axes).
double measure(int cdim, int sdim, const double J[3][3]) {
  if (cdim == sdim) {
    if (cdim == 2) return std::fabs(J[0][0] * J[1][1] - J[0][1] * J[1][0]);
    return std::fabs(J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                     J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                     J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]));
  }
  double G[2][2] = {};
  for (int a = 0; a < cdim; ++a)
    for (int b = 0; b < cdim; ++b)
      for (int d = 0; d < sdim; ++d) G[a][b] += J[d][a] * J[d][b];
  if (cdim == 1) return std::sqrt(G[0][0]);
  return std::sqrt(G[0][0] * G[1][1] - G[0][1] * G[1][0]);
}

// General cells: Q1 (2^cdim vertices, lexicographic) or P1 (cdim + 1 vertices)
// of dimension cdim in a space of dimension sdim.
void cells_fill(bool q1, int cdim, int sdim, const double *vert, int64_t n_cells,
                const int64_t *cells, const Rule &r, const double *nodal, double *x, double *rho,
                double *jxw, double *cen) {
  const int nv = q1 ? 1 << cdim : cdim + 1;
  for (int64_t c = 0; c < n_cells; ++c) {
    double V[8][3] = {};
    double Vr[8];
    for (int i = 0; i < nv; ++i) {
      const int64_t vid = cells[c * nv + i];
      for (int d = 0; d < sdim; ++d) V[i][d] = vert[vid * sdim + d];
      Vr[i] = nodal ? nodal[vid] : 1.0;
    }
    double cc[3] = {0, 0, 0};
    for (int i = 0; i < nv; ++i)
      for (int d = 0; d < sdim; ++d) cc[d] += V[i][d];
    for (int d = 0; d < sdim; ++d) cen[c * sdim + d] = cc[d] / nv;
    for (int q = 0; q < r.nq; ++q) {
      const double *xi = &r.xi[(size_t)q * cdim];
      double phi[8], dphi[8][3] = {};
      if (q1) {
        for (int v = 0; v < nv; ++v) {
          phi[v] = 1.0;
          for (int k = 0; k < cdim; ++k) {
            const double f = ((v >> k) & 1) ? xi[k] : 1.0 - xi[k];
            phi[v] *= f;
          }
          for (int k = 0; k < cdim; ++k) {
            double g = ((v >> k) & 1) ? 1.0 : -1.0;
            for (int m = 0; m < cdim; ++m)
              if (m != k) g *= ((v >> m) & 1) ? xi[m] : 1.0 - xi[m];
            dphi[v][k] = g;
          }
        }
      } else {
        phi[0] = 1.0;
        for (int k = 0; k < cdim; ++k) {
          phi[k + 1] = xi[k];
          phi[0] -= xi[k];
          dphi[0][k] = -1.0;
          dphi[k + 1][k] = 1.0;
        }
      }
      double xp[3] = {0, 0, 0}, rv = 0.0, J[3][3] = {};
      for (int v = 0; v < nv; ++v) {
        for (int d = 0; d < sdim; ++d) {
          xp[d] += V[v][d] * phi[v];
          for (int k = 0; k < cdim; ++k) J[d][k] += V[v][d] * dphi[v][k];
        }
        rv += Vr[v] * phi[v];
      }
      const int64_t g = c * r.nq + q;
      for (int d = 0; d < sdim; ++d) x[g * sdim + d] = xp[d];
      rho[g] = rv;
      jxw[g] = measure(cdim, sdim, J) * r.w[q];
    }
  }
}

double l1_mass(int64_t n, const double *rho, const double *jxw) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += std::fabs(rho[i]) * jxw[i];
  return s;
}

}  // namespace

extern "C" int sdot_mesh_structured(int dim, const int *n_sub, const double *lo,
                                    const double *hi, int qorder, const double *nodal_rho,
                                    int normalize, double *x, double *w, double *centers) {
  if ((dim != 2 && dim != 3) || !n_sub || !lo || !hi || !x || !w || !centers) return SDOT_ERR_INVALID;
  if (qorder < 1 || qorder > 16) return SDOT_ERR_INVALID;
  int ns[3] = {n_sub[0], n_sub[1], dim == 3 ? n_sub[2] : 1};
  for (int d = 0; d < dim; ++d)
    if (ns[d] < 1 || !(hi[d] > lo[d])) return SDOT_ERR_INVALID;
  const Rule r = tensor_rule(dim, qorder);
  const int64_t n_cells = (int64_t)ns[0] * ns[1] * ns[2];
  const int64_t nQ = n_cells * r.nq;
  std::vector<double> rho(nQ), jxw(nQ);
  const int64_t nvert = (int64_t)(ns[0] + 1) * (ns[1] + 1) * (dim == 3 ? ns[2] + 1 : 1);
  std::vector<double> nodal(nvert, 1.0);
  if (nodal_rho) std::memcpy(nodal.data(), nodal_rho, sizeof(double) * nvert);
  structured_fill(dim, ns, lo, hi, r, nodal.data(), x, rho.data(), jxw.data(), centers);
  if (normalize) {
    const double L = l1_mass(nQ, rho.data(), jxw.data());
    if (!(L > 0)) return SDOT_ERR_INVALID;
    for (auto &v : nodal) v /= L;
    structured_fill(dim, ns, lo, hi, r, nodal.data(), x, rho.data(), jxw.data(), centers);
  }
  for (int64_t i = 0; i < nQ; ++i) w[i] = rho[i] * jxw[i];
  return SDOT_OK;
}

extern "C" int sdot_mesh_simplex(int dim, int64_t n_vertices, const double *vertices,
                                 int64_t n_cells, const int64_t *cells, int qorder,
                                 const double *nodal_rho, int normalize, double *x, double *w,
                                 double *centers) {
  if ((dim != 2 && dim != 3) || !vertices || !cells || !x || !w || !centers) return SDOT_ERR_INVALID;
  Rule r;
  if (!simplex_rule(dim, qorder, r)) return SDOT_ERR_UNSUPPORTED;
  for (int64_t i = 0; i < n_cells * (dim + 1); ++i)
    if (cells[i] < 0 || cells[i] >= n_vertices) return SDOT_ERR_INVALID;
  const int64_t nQ = n_cells * r.nq;
  std::vector<double> rho(nQ), jxw(nQ);
  std::vector<double> nodal(n_vertices, 1.0);
  if (nodal_rho) std::memcpy(nodal.data(), nodal_rho, sizeof(double) * n_vertices);
  simplex_fill(dim, vertices, n_cells, cells, r, nodal.data(), x, rho.data(), jxw.data(), centers);
  if (normalize) {
    const double L = l1_mass(nQ, rho.data(), jxw.data());
    if (!(L > 0)) return SDOT_ERR_INVALID;
    for (auto &v : nodal) v /= L;
    simplex_fill(dim, vertices, n_cells, cells, r, nodal.data(), x, rho.data(), jxw.data(),
                 centers);
  }
  for (int64_t i = 0; i < nQ; ++i) w[i] = rho[i] * jxw[i];
  return SDOT_OK;
}

extern "C" int sdot_mesh_cells(int kind, int cell_dim, int spacedim, int64_t n_vertices,
                               const double *vertices, int64_t n_cells, const int64_t *cells,
                               int qorder, const double *nodal_rho, int normalize, double *x,
                               double *w, double *centers) {
  if (kind != SDOT_CELL_SIMPLEX && kind != SDOT_CELL_Q1) return SDOT_ERR_INVALID;
  if (cell_dim < 2 || cell_dim > 3 || spacedim < cell_dim || spacedim > 3) return SDOT_ERR_INVALID;
  if (!vertices || !cells || !x || !w || !centers || n_cells < 1 || n_vertices < 1)
    return SDOT_ERR_INVALID;
  const bool q1 = kind == SDOT_CELL_Q1;
  Rule r;
  if (q1) {
    if (qorder < 1 || qorder > 16) return SDOT_ERR_INVALID;
    r = tensor_rule(cell_dim, qorder);
  } else if (!simplex_rule(cell_dim, qorder, r)) {
    return SDOT_ERR_UNSUPPORTED;
  }
  const int nv = q1 ? 1 << cell_dim : cell_dim + 1;
  for (int64_t i = 0; i < n_cells * nv; ++i)
    if (cells[i] < 0 || cells[i] >= n_vertices) return SDOT_ERR_INVALID;
  const int64_t nQ = n_cells * r.nq;
  std::vector<double> rho(nQ), jxw(nQ);
  std::vector<double> nodal(n_vertices, 1.0);
  if (nodal_rho) std::memcpy(nodal.data(), nodal_rho, sizeof(double) * n_vertices);
  cells_fill(q1, cell_dim, spacedim, vertices, n_cells, cells, r, nodal.data(), x, rho.data(),
             jxw.data(), centers);
  if (normalize) {
    const double L = l1_mass(nQ, rho.data(), jxw.data());
    if (!(L > 0)) return SDOT_ERR_INVALID;
    for (auto &v : nodal) v /= L;
    cells_fill(q1, cell_dim, spacedim, vertices, n_cells, cells, r, nodal.data(), x, rho.data(),
               jxw.data(), centers);
  }
  for (int64_t i = 0; i < nQ; ++i) w[i] = rho[i] * jxw[i];
  return SDOT_OK;
}
