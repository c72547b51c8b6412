/*
 * sdot_oracle.h -- FP64 CPU restatement of SemiDiscreteOT's regularized
 * semi-discrete dual evaluation.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path
 * is compared against, and the "port" CPU baseline that bench.py times.  It
 * is never linked into, loaded by, or called from the product library
 * (semidiscreteot_amd); only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it.
 *
 * Every function cites the reference file:line it restates.  Parity status of
 * each function (what pins it) is listed at the top of sdot_oracle.c.
 */
#ifndef SDOT_ORACLE_H
#define SDOT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Source measure as deal.II's FEValues hands it to local_assemble
 * (SotSolver.cc:788-790): per cell, n_q quadrature points, their density
 * value and JxW, plus the cell center used for the candidate query
 * (SotSolver.cc:801). Arrays are cell-major, q fastest. */
typedef struct {
  int dim;                /* 2 or 3 */
  int64_t n_cells;
  int nq;                 /* quadrature points per cell (uniform) */
  const double *x;        /* [n_cells*nq*dim] quadrature points */
  const double *rho;      /* [n_cells*nq] density at q (FE interpolated) */
  const double *jxw;      /* [n_cells*nq] JxW */
  const double *center;   /* [n_cells*dim] cell centers (vertex average) */
} orc_source_t;

typedef struct {
  int64_t n;
  const double *y;        /* [n*dim] target points */
  const double *nu;       /* [n] target weights */
} orc_target_t;

enum { ORC_THRESH_POINTWISE = 0, ORC_THRESH_INTEGRAL = 1, ORC_THRESH_GEOMETRIC = 2 };
enum { ORC_CTRL_L1NORM = 0, ORC_CTRL_COMPONENTWISE = 1 };
enum { ORC_CONVERGED = 0, ORC_MAX_ITER = 1, ORC_LINESEARCH_FAIL = 2, ORC_NONFINITE = 3 };

/* Subset of SotParameterManager::SolverParameters (ParameterManager.h:56-72)
 * that the evaluation and the BFGS driver read. */
typedef struct {
  double epsilon;
  double tau;
  double tolerance;
  int use_log_sum_exp;
  int threshold_type;
  int control_type;
  int max_iterations;
  int n_threads;          /* 0 = all cores */
  int faithful_copier;    /* 1: per-cell dense CopyData buffers as the reference */
} orc_params_t;

/* Members of SotSolver that persist between evaluate_functional calls
 * (SotSolver.h:505-514). */
typedef struct {
  double covering_radius;     /* R0 (solve() stores 1.1 x raw value) */
  double min_target_density;  /* 0 until solve() sets it */
  double C_global;
  double global_functional;
  double last_threshold;
} orc_state_t;

typedef struct {
  int status;
  int iterations;
  int evaluations;
  double functional;
  double check_value;
} orc_solve_info_t;

/* ---- quadrature / mesh (deal.II stand-ins) ---- */
void orc_gauss_legendre01(int n, double *pts, double *wts);
int orc_structured_source(int dim, const int *n_sub, const double *lo,
                          const double *hi, int qorder,
                          const double *nodal_rho, double *x, double *rho,
                          double *jxw, double *center);
int orc_simplex_source(int dim, int64_t n_vertices, const double *vertices,
                       int64_t n_cells, const int64_t *cells, int qorder,
                       const double *nodal_rho, double *x, double *rho,
                       double *jxw, double *center);
int orc_cells_source(int q1, int cdim, int sdim, int64_t n_vertices,
                     const double *vertices, int64_t n_cells, const int64_t *cells,
                     int qorder, const double *nodal_rho, double *x, double *rho,
                     double *jxw, double *center);
double orc_normalize_density(int64_t n_q, const double *rho, const double *jxw);

/* ---- thresholds ---- */
double orc_covering_radius(const orc_source_t *s, const orc_target_t *t);
double orc_distance_threshold(const orc_params_t *p, const orc_state_t *st,
                              const orc_source_t *s, const orc_target_t *t,
                              const double *psi);

/* ---- evaluation ---- */
int orc_evaluate(const orc_source_t *s, const orc_target_t *t,
                 const orc_params_t *p, orc_state_t *st, const double *psi,
                 double *grad, double *F);
int orc_evaluate_R(const orc_source_t *s, const orc_target_t *t,
                   const orc_params_t *p, double R, const double *psi,
                   double *grad, double *F, double *C);
int orc_candidate_stats(const orc_source_t *s, const orc_target_t *t, double R,
                        int64_t *count, uint64_t *hash);
int orc_discarded_fraction(const orc_source_t *s, const orc_target_t *t,
                           const orc_params_t *p, double R, const double *psi,
                           int64_t n_list, const int64_t *q_list, double *frac);
int orc_hessian(const orc_source_t *s, const orc_target_t *t,
                const orc_params_t *p, double R, const double *psi, double *H);

/* ---- barycenters (Lloyd step) ---- */
/* compute_weighted_barycenters_euclidean (SotSolver.cc:871-971) at threshold R:
 * bary [N*dim]. */
int orc_barycenters_R(const orc_source_t *s, const orc_target_t *t,
                      const orc_params_t *p, double R, const double *psi,
                      double *bary);

/* ---- conditional densities and transport maps ---- */
int orc_conditional_density(const orc_target_t *t, int dim, const double *psi,
                            double epsilon, double R, int64_t n_points,
                            const double *points, const double *density,
                            int64_t n_idx, const int64_t *indices, double *out);
/* strategy 0 = modal, 1 = barycentric */
int orc_transport_map(int strategy, const orc_target_t *t, int dim,
                      const double *psi, double epsilon, double radius,
                      int64_t n_points, const double *points, double *mapped);

/* ---- multilevel potential transfer ---- */
/* SoftmaxRefinement::compute_refinement (SoftmaxRefinement.cc:200-280) with
 * its local_assemble (:61-198).  The coarse-to-fine child lists of the level
 * are CSR: children of coarse point c are child_idx[child_off[c] ..
 * child_off[c+1]) for c < n_parents. */
int orc_softmax_refinement(const orc_source_t *s, int64_t n_fine,
                           const double *y_fine, const orc_target_t *coarse,
                           const double *psi_coarse, int64_t n_parents,
                           const int64_t *child_off, const int64_t *child_idx,
                           double lambda, double threshold, int use_lse,
                           int n_threads, double *psi_fine);

/* ---- BFGS driver ---- */
int orc_solve(const orc_source_t *s, const orc_target_t *t,
              const orc_params_t *p, orc_state_t *st, double *psi,
              orc_solve_info_t *info);

/* textbook objective hooks used to pin the BFGS restatement */
typedef double (*orc_fun_t)(void *ctx, const double *x, double *g);
int orc_bfgs_generic(orc_fun_t fun, void *ctx, int64_t n, double *x,
                     int max_iterations, double tol_l2, orc_solve_info_t *info);

#ifdef __cplusplus
}
#endif
#endif
