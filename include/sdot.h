/*
 * sdot.h -- C ABI of the MI355X (gfx950) regularized semi-discrete OT dual
 * evaluation.  Plain pointers and sizes only; no exceptions cross this
 * boundary (every call returns an sdot_status, details via sdot_last_error).
 *
 * Each entry point names the member of the reference's C++ class
 * SotSolver<dim,spacedim> (include/SemiDiscreteOT/solvers/SotSolver.h) that
 * it replaces.  Host arrays are copied at the call; the library owns every
 * device buffer and frees it in sdot_ctx_destroy.  One context per host
 * thread; calls on a context are ordered on the context's HIP stream.
 */
#ifndef SDOT_H
#define SDOT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sdot_ctx sdot_ctx;

typedef enum {
  SDOT_OK = 0,
  SDOT_ERR_INVALID = -1,   /* bad argument / shape (reference: AssertThrow, runtime_error) */
  SDOT_ERR_STATE = -2,     /* source/target not set (SotSolver.cc:110-112) */
  SDOT_ERR_HIP = -3,       /* HIP runtime failure */
  SDOT_ERR_NOMEM = -4,
  SDOT_ERR_PARSE = -5,     /* .prm parse error */
  SDOT_ERR_UNSUPPORTED = -6
} sdot_status;

/* Outcome of a solve (the reference keeps non-convergence out of the
 * exception path: SotSolver.cc:207-211). */
typedef enum {
  SDOT_SOLVE_CONVERGED = 0,
  SDOT_SOLVE_MAX_ITER = 1,
  SDOT_SOLVE_LINESEARCH_FAIL = 2,
  SDOT_SOLVE_NONFINITE = 3
} sdot_solve_status;

typedef enum { SDOT_THRESH_POINTWISE = 0, SDOT_THRESH_INTEGRAL = 1, SDOT_THRESH_GEOMETRIC = 2 } sdot_threshold_type;
typedef enum { SDOT_CTRL_L1NORM = 0, SDOT_CTRL_COMPONENTWISE = 1 } sdot_control_type;

/* Arithmetic of the pair kernels.
 *  FP64:     everything in double (parity mode, matches the FP64 CPU path).
 *  FP32_LSE: exponent argument and max in double, exp2 in FP32, sums and
 *            every reduction in double.  Plain (non-LSE) mode always runs FP64. */
typedef enum { SDOT_PREC_FP64 = 0, SDOT_PREC_FP32_LSE = 1 } sdot_precision;

/* Line search of the BFGS driver.
 *  REFERENCE:   deal.II's line minimisation as SotSolver::solve uses it.
 *  NOISE_AWARE: the same search until the decrease the first trial predicts
 *               falls below 100x the rounding noise of the functional (about
 *               2e-15 |F| in FP64, 2e-9 |F| in FP32_LSE); from then on the
 *               functional differences it compares are trapezoid integrals of
 *               the measured slopes (Hager & Zhang's approximate Wolfe idea)
 *               and the first trial step is 1.  Lets a solve reach tolerances
 *               at which F differences are below their rounding noise. */
typedef enum { SDOT_LS_REFERENCE = 0, SDOT_LS_NOISE_AWARE = 1 } sdot_line_search;

/* Initial inverse Hessian of the L-BFGS two-loop recursion.
 *  NONE:     the identity, as deal.II's SolverBFGS (SotSolver.cc:190-198).
 *  DIAGONAL: diag(epsilon / nu_j), the inverse of the Hessian's diagonal
 *            m_j / epsilon at the solution (m = nu there).  Its first step is
 *            the target-side Sinkhorn update psi_j -= epsilon (m_j / nu_j - 1)
 *            to first order; it makes the iteration independent of the scale
 *            of the target weights. */
typedef enum { SDOT_PRECOND_NONE = 0, SDOT_PRECOND_DIAGONAL = 1 } sdot_preconditioner;

/* SotParameterManager::SolverParameters (ParameterManager.h:56-72), the keys
 * of subsection rsot_solver (ParameterManager.cc:64-95), plus `precision`. */
typedef struct {
  uint32_t max_iterations;        /* 1000 */
  double tolerance;               /* 1 */
  int32_t solver_control_type;    /* sdot_control_type, "l1norm" */
  double epsilon;                 /* 1e-3 */
  double tau;                     /* 1e-8 */
  int32_t distance_threshold_type;/* sdot_threshold_type, "pointwise" */
  int32_t verbose_output;         /* 1 */
  int32_t solver_type;            /* 0 = BFGS (the only type) */
  uint32_t quadrature_order;      /* 3 */
  uint32_t nb_points;             /* 1000 */
  uint32_t number_of_threads;     /* 0 */
  int32_t use_epsilon_scaling;    /* 0 */
  double epsilon_scaling_factor;  /* 2.0 */
  uint32_t epsilon_scaling_steps; /* 5 */
  int32_t use_log_sum_exp_trick;  /* 0 */
  int32_t precision;              /* sdot_precision, FP64 */
  int32_t line_search;            /* sdot_line_search, REFERENCE */
  int32_t preconditioner;         /* sdot_preconditioner, NONE */
  /* Value bound of the pair sums (this library's own; no counterpart in the
   * reference).  Bins of candidate pairs whose terms all lie below
   * 2^-prune_bits of their q-point's largest weighted term are skipped: a
   * q-point's sum loses less than N 2^-prune_bits of itself and a target's
   * mass less than 2^-prune_bits.  < 0: the precision's default: FP64 off
   * (every candidate pair, as the reference sums them); FP32_LSE 50 bits,
   * below the FP32 exponentials' own rounding, where it pays (at least 2^32
   * pairs, and sqrt(100 ln2 epsilon) at most a fifth of the diagonal of the
   * source box and half the truncation radius), else off.  0: off.  Plain (non-LSE) mode and
   * targets with a zero weight turn it off. */
  double prune_bits;              /* -1 */
} sdot_params;

typedef struct {
  int32_t status;          /* sdot_solve_status */
  int32_t iterations;      /* SotSolver::get_last_iteration_count */
  int32_t evaluations;     /* functional evaluations incl. line search */
  double functional;       /* SotSolver::get_last_functional_value */
  double check_value;      /* last stopping-test value */
  double seconds;
} sdot_solve_info;

typedef struct {
  double last_functional;        /* get_last_functional_value (SotSolver.h:240) */
  double C_global;               /* get_C_global (SotSolver.h:256) */
  double last_distance_threshold;/* get_last_distance_threshold (SotSolver.h:252) */
  double covering_radius;        /* R0 x 1.1 after a solve (SotSolver.cc:132) */
  double min_target_density;     /* SotSolver.cc:133 */
  int32_t last_iteration_count;  /* get_last_iteration_count (SotSolver.cc:572-576) */
  int32_t converged;             /* get_convergence_status (SotSolver.cc:578-582) */
  int64_t last_pair_count;       /* (q, candidate) pairs of the last evaluation, when counted */
} sdot_state;

/* Kernel timing gathered with HIP events on the context stream.  The dense
 * Hessian (sdot_hessian) adds its P-build + SYRK chunks, from the first chunk's
 * start to the last one's end, to ms_hessian and counts the call. */
typedef struct {
  int64_t evaluations;
  double ms_lse_max, ms_lse_sum, ms_qfin, ms_grad, ms_total;
  int64_t hessians;
  double ms_hessian;
} sdot_profile;

/* ---- library / context ---------------------------------------------- */
const char *sdot_version(void);
/* Thread-local message of the last failing call on this thread. */
const char *sdot_last_error(void);
/* SotSolver::SotSolver(comm) (SotSolver.cc:3-20).  device = HIP ordinal. */
int sdot_ctx_create(int device, sdot_ctx **out);
int sdot_ctx_destroy(sdot_ctx *ctx);

/* ---- parameters --------------------------------------------------------- */
/* Defaults of SolverParameters (ParameterManager.h:56-72). */
int sdot_params_default(sdot_params *p);
/* Parse `subsection rsot_solver ... end` of a deal.II .prm text
 * (ParameterManager.cc:64-95); keys outside it are ignored. */
int sdot_params_parse_prm(const char *text, sdot_params *p);
int sdot_params_load_prm(const char *path, sdot_params *p);
/* The rest of a SotParameterManager parameter file (ParameterManager.cc:
 * 19-146, ParameterManager.h:42-121) and the `lloyd parameters` subsection
 * of LloydParameterManager (ParameterManager.cc:412-422, ParameterManager.h:
 * 202-205): everything the reference's run/ parameter files set.  Strings are
 * NUL-terminated; a value longer than its buffer is a parse error. */
typedef struct {
  uint32_t n_refinements;              /* "number of refinements" (0) */
  char grid_generator_function[64];    /* "grid generator function" */
  char grid_generator_arguments[256];  /* "grid generator arguments" */
  int32_t use_tetrahedral_mesh;        /* "use tetrahedral mesh" (false) */
  int32_t use_custom_density;          /* "use custom density" (false) */
  char density_file_path[256];         /* "density file path" */
  char density_file_format[16];        /* "density file format" ("vtk") */
  char density_field_name[64];         /* "density field name" ("normalized_density") */
} sdot_mesh_params;

/* SotParameterManager::MultilevelParameters (ParameterManager.h:78-98). */
typedef struct {
  int32_t source_enabled;              /* false */
  int32_t source_min_vertices;         /* 1000 */
  int32_t source_max_vertices;         /* 10000 */
  char source_hierarchy_dir[256];      /* "output/data_multilevel/source_multilevel" */
  int32_t target_enabled;              /* false */
  int32_t target_min_points;           /* 100 */
  int32_t target_max_points;           /* 1000 */
  char target_hierarchy_dir[256];      /* "output/data_multilevel/target_multilevel" */
  int32_t use_softmax_potential_transfer; /* true */
  int32_t use_python_clustering;       /* false */
  char python_script_name[128];        /* "multilevel_clustering_scipy.py" */
  char output_prefix[256];             /* "output/multilevel/sot" */
} sdot_multilevel_params;

#define SDOT_MAX_DENSITY_INDICES 256
typedef struct {
  sdot_params solver;                  /* subsection rsot_solver */
  char selected_task[64];              /* "selected_task" ("") */
  int32_t io_coding;                   /* "io_coding": 0 = txt (default), 1 = bin */
  sdot_mesh_params source, target;     /* subsection mesh_generation / source, target */
  sdot_multilevel_params multilevel;   /* subsection multilevel_parameters */
  char power_diagram_implementation[16]; /* power_diagram_parameters ("dealii") */
  double transport_map_truncation_radius; /* transport_map_parameters (-1 = disabled) */
  int32_t n_density_indices;           /* conditional_density_parameters "indices" */
  uint32_t density_indices[SDOT_MAX_DENSITY_INDICES];
  char density_potential_folder[256];  /* "potential_folder" ("") */
  double density_truncation_radius;    /* "truncation_radius" (-1 = disabled) */
  uint32_t lloyd_max_iterations;       /* lloyd parameters "max iterations" (1000) */
  double lloyd_relative_tolerance;     /* "relative tolerance" (1e-8) */
} sdot_prm_file;

/* Defaults of every field (the declarations cited above). */
int sdot_prm_file_default(sdot_prm_file *f);
/* Parse a whole parameter file: the subsections above directly under
 * `subsection SotParameterManager` (or at the top level), rsot_solver as
 * sdot_params_parse_prm reads it; keys these subsections do not declare are
 * a parse error (deal.II rejects them), other subsections are skipped. */
int sdot_prm_file_parse(const char *text, sdot_prm_file *f);
int sdot_prm_file_load(const char *path, sdot_prm_file *f);

/* SotSolver::configure (SotSolver.cc:63-70). */
int sdot_set_params(sdot_ctx *ctx, const sdot_params *p);

/* ---- measures ----------------------------------------------------------- */
/* SotSolver::setup_source (SotSolver.h:186-190, SotSolver.cc:44-53): the
 * values FEValues produces per locally owned cell, pulled once.
 *   x       [n_cells*nq*dim]  quadrature points, cell-major, q fastest
 *   w       [n_cells*nq]      density(x_q) * JxW_q
 *   centers [n_cells*dim]     cell centers (candidate query points, SotSolver.cc:801) */
int sdot_set_source(sdot_ctx *ctx, int dim, int64_t n_cells, int nq,
                    const double *x, const double *w, const double *centers);
/* SotSolver::setup_target (SotSolver.h:197-198, SotSolver.cc:55-61). */
int sdot_set_target(sdot_ctx *ctx, int64_t n, const double *y, const double *nu);
/* No counterpart in the reference (audit, like sdot_candidate_stats): the
 * target layout sdot_set_target builds on the device.  pos_r / pos_m: N
 * positions of each target (caller order) in row-major bin order and in Morton
 * order; tstart: n_bins + 1 bin offsets; grid: origin[3], bin width, bins per
 * axis[3].  Any output may be NULL. */
int sdot_target_layout(sdot_ctx *ctx, int32_t *pos_r, int32_t *pos_m, int32_t *tstart,
                       int64_t *n_bins, double *grid);

/* ---- the hot path --------------------------------------------------------- */
/* SotSolver::evaluate_functional (SotSolver.h:335-336, SotSolver.cc:217-377):
 * threshold from the stored state, assembly, linear term.  grad may be NULL. */
int sdot_eval(sdot_ctx *ctx, const double *psi, double *grad, double *F, double *C_global);
/* Same with an explicit distance threshold R (no state update). */
int sdot_eval_R(sdot_ctx *ctx, double R, const double *psi, double *grad, double *F,
                double *C_global);
/* Device-pointer variant for callers whose psi/grad already live in HBM:
 * d_psi, d_grad: N doubles in original target order; d_FC: 2 doubles (F, C).
 * Asynchronous on the context stream (sdot_synchronize to wait). */
int sdot_eval_device(sdot_ctx *ctx, double R, const double *d_psi, double *d_grad, double *d_FC);
/* No counterpart in the reference: sdot_eval_device for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results; it calls no
 * hipMalloc, hipFree or other allocating routine.  Every per-evaluation
 * buffer comes from one device array of 128 MiB per GPU (held until the
 * process ends) shared by every call on that GPU, from every context: calls
 * that overlap on one GPU race.  The FP32_LSE sweep runs unseeded; the segment
 * counts shrink until the buffers fit, and an evaluation that does not fit
 * with one segment per sweep (about 160 N + 96 nQ bytes) returns
 * SDOT_ERR_INVALID.  Nothing in the array has to start at zero. */
int sdot_eval_device_noalloc(sdot_ctx *ctx, double R, const double *d_psi, double *d_grad,
                             double *d_FC);
int sdot_synchronize(sdot_ctx *ctx);

/* SotSolver::evaluate_weighted_barycenters / compute_weighted_barycenters_euclidean
 * (SotSolver.cc:585-610, 871-971): per target j, sum over q of (w_q/S_q) nu_j e_qj x_q
 * divided by nu_j (the first moment of the transport plan's column j over nu_j),
 * with the threshold from the stored state (the functional state is not changed).
 * bary [N*dim] out, caller order.  The _R form takes the threshold explicitly. */
int sdot_barycenters(sdot_ctx *ctx, const double *psi, double *bary);
int sdot_barycenters_R(sdot_ctx *ctx, double R, const double *psi, double *bary);

/* SotSolver::compute_hessian (declared at SotSolver.h:338-344, no reference
 * definition): dense H (row-major, leading dimension ld >= N) of the
 * functional at psi with the current threshold rule.  N <= 65536
 * (SDOT_ERR_UNSUPPORTED above). */
int sdot_hessian(sdot_ctx *ctx, const double *psi, double *H, int64_t ld);

/* SotSolver::solve (SotSolver.h:211-212, SotSolver.cc:105-215): L-BFGS with
 * the reference's line search and stopping test; psi in/out (size N).  In
 * FP32_LSE precision the iteration runs on FP32 evaluations until its stopping
 * value is within 2x the tolerance (or a line search fails), then continues
 * from there on FP64 evaluations within the same iteration budget, so that
 * the stopping decision is taken on an FP64 gradient.  From N = 32768 on, the
 * L-BFGS vectors stay in device memory for the length of the call (18 N
 * doubles); the environment variable SDOT_LBFGS_DEVICE_MIN moves that length. */
int sdot_solve(sdot_ctx *ctx, double *psi, sdot_solve_info *info);

/* ---- multilevel potential transfer ---------------------------------------- */
/* SoftmaxRefinement<dim>::compute_refinement (SoftmaxRefinement.h, SoftmaxRefinement.cc:
 * 200-280, with local_assemble :61-198) over the context's source measure (the
 * source the reference constructs it with, SemiDiscreteOT.cc:776-784): the
 * potential of a fine target level from the coarse level's,
 *   psi_f = -lambda log sum_q (w_q / Z_q) exp(-|x_q - y_f|^2 / (2 lambda)),
 *   Z_q   = sum_c nu_c exp((psi_c - |x_q - y_c|^2 / 2) / lambda),
 * c over the coarse points within distance_threshold of the center of the cell
 * of q and of x_q, f over the children of those coarse points within
 * distance_threshold of x_q; fine points that receive nothing get 0.
 * use_log_sum_exp_trick selects the max-shifted form of the reference.
 *   y_fine [n_fine*dim]; y_coarse [n_coarse*dim], nu_coarse, psi_coarse [n_coarse]
 *   child lists of the level (child_indices[level] of the reference) as CSR:
 *   the children of coarse point c < n_parents are
 *   children[child_offsets[c] .. child_offsets[c+1]) (child_offsets[0] = 0);
 *   coarse points c >= n_parents have none and children outside [0, n_fine)
 *   are skipped, as the reference does after reporting them.
 * The fine target density of the reference's signature is not read by its
 * algorithm and is not taken.  The context's target measure is left as it is.
 * FP64; psi_fine [n_fine] out. */
int sdot_softmax_refinement(sdot_ctx *ctx, int64_t n_fine, const double *y_fine,
                            int64_t n_coarse, const double *y_coarse, const double *nu_coarse,
                            const double *psi_coarse, int64_t n_parents,
                            const int64_t *child_offsets, const int64_t *children,
                            double lambda, double distance_threshold,
                            int use_log_sum_exp_trick, double *psi_fine);

/* ---- point queries after a solve ------------------------------------------ */
/* SotSolver::get_potential_conditioned_density (SotSolver.cc:973-1046) over the
 * context's target measure at epsilon = the configured epsilon (the reference's
 * current_epsilon) and threshold R (its current_distance_threshold; R >= DBL_MAX:
 * every target): at each point p (the reference's DoF support points) the
 * targets j with |p - y_j| <= R,
 *   c_i(p) = rho_p exp(e_pi - M_p) / sum_j nu_j exp(e_pj - M_p),
 *   e_pj = (psi_j - |p - y_j|^2 / 2) / eps,  M_p = max_j e_pj,
 * 0 where i is not among p's candidates or the sum is not positive.
 *   points [n_points*dim], density [n_points] (rho at the points),
 *   indices [n_idx] (potential_indices), out [n_idx*n_points] row k = index k. */
int sdot_conditional_density(sdot_ctx *ctx, const double *psi, double R, int64_t n_points,
                             const double *points, const double *density, int64_t n_idx,
                             const int64_t *indices, double *out);

/* ModalStrategy / BarycentricStrategy::compute_map (OptimalTransportPlan.cc:
 * 99-189, 218-310) over the context's target measure: at each point x the
 * targets with |x - y_j| < radius (every target when radius <= 0; the nearest
 * target when none is that close),
 *   MODAL:       y_j of the maximum of psi_j - |x - y_j|^2 / 2 + epsilon log nu_j
 *                (first maximum in index order),
 *   BARYCENTRIC: sum_j nu_j e^(e_j - M) y_j / sum_j nu_j e^(e_j - M),
 *                e_j = (psi_j - |x - y_j|^2 / 2) / epsilon (the nearest target
 *                where the sum is not positive).
 * Nearest: lowest index among equal distances.  mapped [n_points*dim] out. */
typedef enum { SDOT_MAP_MODAL = 0, SDOT_MAP_BARYCENTRIC = 1 } sdot_map_strategy;
int sdot_transport_map(sdot_ctx *ctx, int strategy, const double *psi, double epsilon,
                       double radius, int64_t n_points, const double *points, double *mapped);

/* ---- thresholds and audit ------------------------------------------------ */
/* compute_covering_radius (SotSolver.cc:431-473), raw (without the 1.1). */
int sdot_covering_radius(sdot_ctx *ctx, double *r0);
/* compute_distance_threshold (SotSolver.cc:379-429) for psi with the state. */
int sdot_distance_threshold(sdot_ctx *ctx, const double *psi, double *R);
/* set_distance_threshold (SotSolver.h:262). */
int sdot_set_distance_threshold(sdot_ctx *ctx, double R);
/* Candidate set of every cell under threshold R (find_nearest_target_points,
 * SotSolver.cc:553-570), as count and order-free hash per cell (original
 * cell order): hash = sum over members j of mix(j+1), mix(v) = v*0x9E3779B97F4A7C15,
 * v ^= v>>29.  Returns the total pair count in *pairs (may be NULL). */
int sdot_candidate_stats(sdot_ctx *ctx, double R, int64_t *count, uint64_t *hash,
                         int64_t *pairs);
/* Audit of a truncation radius: for every quadrature point (caller order,
 * cell-major) the log of its candidate partition sum, in FP64,
 *   logZ_q = log sum_{j in C(cell(q))} nu_j exp((psi_j - |x_q - y_j|^2/2)/eps) + |x_q'|^2/(2 eps)
 * (the constant shift is the same for every R, -inf when the sum vanishes).  The mass
 * a radius R discards at q is 1 - exp(logZ_q(R) - logZ_q(inf)). */
int sdot_q_log_partition(sdot_ctx *ctx, double R, const double *psi, double *logZ);
int sdot_get_state(sdot_ctx *ctx, sdot_state *st);
int sdot_set_state(sdot_ctx *ctx, const sdot_state *st);
int sdot_profile_enable(sdot_ctx *ctx, int on);
int sdot_profile_get(sdot_ctx *ctx, sdot_profile *out);

/* ---- batched serving ------------------------------------------------------
 * No counterpart in the reference, whose SotSolver::solve (SotSolver.cc:105-215)
 * runs one solve at a time; each solve here follows that function's algorithm
 * (L-BFGS m = 5 with the same line search and stopping tests).
 *
 * A server borrows the source measure of a context (the resident mesh; that
 * context's sdot_set_source and sdot_ctx_destroy fail while the server
 * exists) and advances many solves against it together: every step runs one
 * evaluation of each active solve through work-list kernels (replayed as one
 * captured hipGraph), with each solve's potential, gradient, L-BFGS history
 * and partial sums resident in HBM.  Solves are admitted in submission order
 * while a slot is free and their buffers fit the HBM budget.  A solve's
 * arithmetic does not depend on which other solves share its steps: its
 * result is bitwise the same alone and in any mix.
 * submit / poll may be called from any thread; step from one thread at a time. */
typedef struct sdot_server sdot_server;

typedef struct {
  int32_t precision;         /* sdot_precision of the pair kernels (default FP32_LSE) */
  int32_t max_slots;         /* solves advanced together (default 256, at most 4096) */
  int64_t hbm_budget_bytes;  /* device arena for per-solve buffers (default 16 GiB) */
  int32_t use_graph;         /* 1: replay one captured hipGraph per step (default 1) */
  int32_t pass_targets;      /* targets per q-major work item (0: tuned or default) */
  int32_t grad_qpoints;      /* q-points per target-major work item (0: tuned or default) */
  double polish_factor;      /* FP32_LSE: a solve switches to FP64 evaluations once its
                                stopping value is within this factor of the tolerance (or
                                an FP32 line search fails), so that it stops on an FP64
                                evaluation; 0: stay FP32 (default 2) */
} sdot_server_config;

typedef struct {
  int64_t n;                   /* number of targets (>= 1) */
  const double *y;             /* [n*dim] target points */
  const double *nu;            /* [n] target weights (> 0) */
  const double *psi0;          /* [n] initial potential, or NULL for zeros */
  double epsilon;              /* regularisation (> 0) */
  double tolerance;            /* stopping tolerance (SotSolver.cc:156-162) */
  int32_t solver_control_type; /* sdot_control_type */
  uint32_t max_iterations;
  double radius;               /* truncation radius R of the candidate sets; <= 0 or
                                  DBL_MAX: every target is a candidate */
  int32_t line_search;         /* sdot_line_search */
  /* epsilon scaling as SemiDiscreteOT::run_sot applies it (SemiDiscreteOT.cc:877-918):
   * solves at epsilon * factor^(steps-1-i), i = 0..steps-1, each to the tolerance,
   * the potential carried over; steps 0 or 1: a single solve at epsilon;
   * factor <= 1 is replaced by 2 (EpsilonScalingHandler.cc:18-27) */
  double epsilon_scaling_factor;
  uint32_t epsilon_scaling_steps;
  int32_t preconditioner;      /* sdot_preconditioner */
} sdot_request;

typedef struct {
  int64_t id;
  int32_t status;              /* sdot_solve_status */
  int32_t iterations, evaluations;
  double functional;           /* F at the final potential */
  double check_value;          /* last stopping-test value */
  double rel_marginal_error;   /* max_j |g_j| / nu_j at the final potential */
  double t_submit, t_admit, t_done;  /* seconds on the server clock */
} sdot_result;

int sdot_server_config_default(sdot_server_config *cfg);
int sdot_server_create(sdot_ctx *source, const sdot_server_config *cfg, sdot_server **out);
int sdot_server_destroy(sdot_server *srv);
/* Copies the request's arrays and builds its target layout; *id >= 0. */
int sdot_server_submit(sdot_server *srv, const sdot_request *req, int64_t *id);
/* Admits what fits, runs one batched step, advances every active solve's
 * line search on the host.  Outputs (may be NULL): solves active after the
 * step, requests still queued. */
int sdot_server_step(sdot_server *srv, int32_t *n_active, int32_t *n_queued);
/* *done = 1 when solve `id` has finished: fills *res and, if psi is not NULL,
 * the final potential (n doubles), then forgets the solve. */
int sdot_server_poll(sdot_server *srv, int64_t id, int32_t *done, sdot_result *res, double *psi);
/* Seconds since the server was created (the clock of sdot_result). */
int sdot_server_clock(sdot_server *srv, double *t);
/* Device bytes the arena would hold for a request of n targets. */
int sdot_server_request_bytes(sdot_server *srv, int64_t n, double radius, int64_t *bytes);
/* Times candidate work-item sizes on synthetic target sets of the server's
 * source and keeps the fastest (pass_targets, grad_qpoints); call before the
 * first submit.  Outputs may be NULL. */
int sdot_server_autotune(sdot_server *srv, int32_t *pass_targets, int32_t *grad_qpoints);

/* ---- source-measure extraction (the FEValues stand-in) --------------------- */
/* Structured Q1 box mesh (GridGenerator::subdivided_hyper_rectangle), QGauss
 * of `qorder` points per direction, Q1-interpolated nodal density (vertices
 * lexicographic, x fastest; NULL = 1).  Outputs sized n_cells*nq*dim etc.,
 * with w = rho*JxW.  If normalize, the nodal density is divided by
 * sum |rho_q| JxW_q first (normalize_density, SemiDiscreteOT.cc:422-449). */
int sdot_mesh_structured(int dim, const int *n_sub, const double *lo, const double *hi,
                         int qorder, const double *nodal_rho, int normalize,
                         double *x, double *w, double *centers);
/* P1 simplices (triangles/tets) with QGaussSimplex(qorder), qorder 1 or 2. */
int sdot_mesh_simplex(int dim, int64_t n_vertices, const double *vertices, int64_t n_cells,
                      const int64_t *cells, int qorder, const double *nodal_rho,
                      int normalize, double *x, double *w, double *centers);

/* General unstructured cells, as GridIn::read_msh / read_vtk meshes reach
 * setup_source (MeshManager.templates.h:80-140): kind SDOT_CELL_Q1 (FE_Q(1),
 * MappingQ1, QGauss(qorder); 2^cell_dim vertices per cell in deal.II's
 * lexicographic order) or SDOT_CELL_SIMPLEX (FE_SimplexP(1),
 * QGaussSimplex(qorder), orders 1-2; cell_dim + 1 vertices), cells of
 * dimension cell_dim in a space of dimension spacedim >= cell_dim (surfaces
 * in 3D: SotSolver<2, 3>; JxW is then the Gram measure).  vertices
 * [n_vertices*spacedim], cells [n_cells*vertices per cell], nodal_rho
 * [n_vertices] or NULL; x [n_cells*nq*spacedim], w [n_cells*nq], centers
 * [n_cells*spacedim] out (nq = qorder^cell_dim for Q1). */
typedef enum { SDOT_CELL_SIMPLEX = 0, SDOT_CELL_Q1 = 1 } sdot_cell_kind;
int sdot_mesh_cells(int kind, int cell_dim, int spacedim, int64_t n_vertices,
                    const double *vertices, int64_t n_cells, const int64_t *cells, int qorder,
                    const double *nodal_rho, int normalize, double *x, double *w,
                    double *centers);

#ifdef __cplusplus
}
#endif
#endif
