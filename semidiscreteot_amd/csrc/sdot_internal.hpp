// sdot_internal.hpp -- context layout and kernel-launcher interface shared by
// the host-side C++ (sdot_api.cpp) and the HIP kernels (sdot_kernels.hip).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sdot.h"

namespace sdot {

// Doubles per target record and per quadrature record (64 bytes: one scalar
// s_load_dwordx16 per record).
constexpr int kRec = 8;
constexpr int kWave = 64;

enum Mode : int { kF64Lse = 0, kF64Plain = 1, kF32Lse = 2 };

// Uniform cubic bin grid. bin_of() is the single binning rule used by the
// host (to sort points into bins) and by the device (to turn query bounds into
// bin ranges). It is monotone non-decreasing in v, which is what makes the
// device's conservative range query never miss a binned point.
struct Grid {
  double o[3];
  double h;
  double inv_h;
  int n[3];
};

#if defined(__HIPCC__)
#define SDOT_HD __host__ __device__
#else
#define SDOT_HD
#endif

SDOT_HD inline int bin_of(double v, double o, double inv_h, int n) {
  double f = floor((v - o) * inv_h);
  if (!(f >= 0.0)) return 0;
  if (f >= (double)n) return n - 1;
  return (int)f;
}

// Rows (x-runs of bins) of a grid.
SDOT_HD inline int64_t grid_rows(const Grid &g) { return (int64_t)g.n[1] * g.n[2]; }

// Row-major index of the bin holding p (the host and the device bin the same
// way).
SDOT_HD inline int bin_linear(const Grid &g, const double *p, int dim) {
  int b[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) b[d] = bin_of(p[d], g.o[d], g.inv_h, g.n[d]);
  return (b[2] * g.n[1] + b[1]) * g.n[0] + b[0];
}

// Morton key of p in the box [lo, hi]: 21 bits per axis in 3D, 31 in 2D.
SDOT_HD inline uint64_t morton_key(const double *p, const double *lo, const double *hi, int dim) {
  const int bits = dim == 3 ? 21 : 31;
  const double scale = (double)((1u << (bits - 1)) - 1) * 2.0;
  uint64_t key = 0;
  uint64_t q[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) {
    const double ext = hi[d] - lo[d];
    double u = ext > 0 ? (p[d] - lo[d]) / ext : 0.0;
    u = u > 0.0 ? u : 0.0;
    u = u < 1.0 ? u : 1.0;
    q[d] = (uint64_t)(u * scale);
  }
  for (int b = bits - 1; b >= 0; --b)
    for (int d = dim - 1; d >= 0; --d) key = (key << 1) | ((q[d] >> b) & 1u);
  return key;
}

// Device buffers and host metadata of one solver context.
struct Dev {
  // ---- source (static after sdot_set_source) ----
  double *qx_m = nullptr;      // [nQ*dim] shifted quadrature points, Morton cell order
  double *qxr_m = nullptr;     // [nQ*dim] raw quadrature points, Morton cell order
  double *qw_m = nullptr;      // [nQ] rho*JxW
  double *qx2_m = nullptr;     // [nQ] |x'|^2 (plain-mode shift)
  double *cc_m = nullptr;      // [n_cells*dim] raw centers, Morton order
  double *blkbox = nullptr;    // [n_blocks*2*dim] AABB of the block's cell centers
  double *blkbox_s = nullptr;  // the same for the small grouping
  double *qbox_s = nullptr;    // [n_blocks_s*2*dim] AABB of the block's q-points, small grouping
  double *cc_r = nullptr;      // [n_cells*4] raw centers, row (bin) order
  int32_t *cstart = nullptr;   // [n_cbins+1] row-order cell offsets per bin
  int64_t *cell_r2m = nullptr; // [n_cells] row-order cell -> Morton cell
  int64_t *cell_m2o = nullptr; // [n_cells] Morton cell -> caller's cell index
  // ---- target (static after sdot_set_target) ----
  double *ty_o = nullptr;      // [N*dim] raw target points, caller order
  double *tnu_o = nullptr;     // [N]
  int32_t *tpos_r = nullptr;   // [N] caller index -> row-order position
  int32_t *tpos_m = nullptr;   // [N] caller index -> Morton position
  int32_t *tr2o = nullptr;     // [N] row-order position -> caller index
  int32_t *tm2o = nullptr;     // [N] Morton position -> caller index
  int32_t *tstart = nullptr;   // [n_tbins+1]
  double *twbox = nullptr;     // [n_twaves*2*dim]
  double *ty_m = nullptr;      // [N*4] raw points, Morton order (covering radius)
  // ---- per-evaluation ----
  double *trec_r = nullptr;    // [N*kRec] row order
  double *trec_m = nullptr;    // [N*kRec] Morton order
  double *qrec_r = nullptr;    // [nQ*kRec] row cell order
  double *qchunk = nullptr;    // [nQ*(dim+2)] GArgs::qchunk
  double *Mpart = nullptr;     // [K1*nQ]
  double *Spart = nullptr;     // [K1*nQ]
  double *Mglob = nullptr;     // [nQ] FP64 LSE: the maximum over the segments
  double *Tpart = nullptr;     // [K1*nQ] FP32: segment maxima of the exponent
  double *Gpart = nullptr;     // [K2*N]
  double *Fpart = nullptr;     // [nFpart]
  double *Cpart = nullptr;     // [nFpart]
  double *PNpart = nullptr;    // [nPpart]
  double *psi = nullptr;       // [N] staging
  double *grad = nullptr;      // [N] staging
  double *out = nullptr;       // [4] F, C, count of q whose FP32 sum lost range
  double *Upart = nullptr;     // [nFpart]
  double *psi_prev = nullptr;  // [N] psi of the previous FP32 evaluation
  double *dpsi_r = nullptr;    // [N] exponent change since then, row order
  double *dpsi_cx = nullptr;   // [ceil(N/64)] its maximum per chunk of 64 records
  double *dpsi_cn = nullptr;   // [ceil(N/64)] its minimum
  double *ucount = nullptr;    // [1] q-points whose FP32 sum lost its range
  // value bound (per evaluation)
  double *psit_r = nullptr;    // [N] psi + eps ln nu, row order
  double *pbin = nullptr;      // [n_tbins] its largest per target bin
  double *prow = nullptr;      // [n_trows] ... per bin row
  double *vcell = nullptr;     // [n_cells] smallest V_q per cell, row order
  double *vbin = nullptr;      // [n_cbins] ... per cell bin
  double *vrow = nullptr;      // [n_crows] ... per cell-bin row
  double *scratch = nullptr;   // misc
  size_t scratch_bytes = 0;
  int64_t *i64scratch = nullptr;
  size_t i64scratch_n = 0;
};

// Quadrature points per lane of the q-major passes (G), lanes per cell (gpc)
// and the resulting 64-lane blocks.
struct Grouping {
  int G = 1, gpc = 1;
  int64_t n_groups = 0, n_blocks = 0;
};

// Everything a kernel launch needs, by value.
struct Plan {
  int dim = 0, nq = 0, G = 1, gpc = 1, mode = kF64Lse;
  int64_t n_cells = 0, nQ = 0, n_groups = 0, n_blocks = 0;
  // big: most points per lane (dense sweeps: fewest record loads per pair);
  // small: at most 4 per lane (truncated sweeps: half-size blocks of cells,
  // fewer lanes idle at the edge of the candidate balls); `small` selects it
  Grouping grp_big, grp_small;
  int small = 0;
  int64_t N = 0, n_twaves = 0;
  int K1 = 1, K2 = 1;
  int64_t nFpart = 0, nPpart = 0;
  Grid tg{}, cg{};
  double origin[3] = {0, 0, 0};
  double eps = 1.0;
  double smax = 0.0;   // candidate rule: s <= smax  <=>  sqrt(s) <= R
  double Rq = 0.0;     // conservative enumeration radius
  double s_in = 0.0;    // inner runs: every pair closer than sqrt(s_in) passes the rule
  int hist = 0;        // FP32 mode: seed the online sweep with the previous evaluation
  int hist_same = 0;   // ... which ran at this radius
  int dense = 0;       // every pair is a candidate (no enumeration, no test)
  double seed_bias = 0.0;  // test hook: raises the FP32 sweep's seeded reference values
  // value bound of the pair sums (VBound): 2^-prune_bits, 0 = off
  double prune_bits = 0.0;
  double nu_min = 0.0;     // smallest target weight (0: some weight is 0, no bound)
  double rq = 0.0;         // largest distance of a cell's q-points from its center
  int64_t n_tbins = 0, n_cbins = 0;  // bins of the target / cell grids
  // the q-major sweeps stream the target records through LDS (record sets
  // larger than one XCD's L2)
};


// Argument blocks of the evaluation passes (one per solve; shared by the
// single-solve grids and the serving work lists).
struct PrepArgs {
  const double *y;     // [N*dim] caller order
  const double *nu;    // [N]
  const double *psi;   // [N]
  const int32_t *pos_r, *pos_m;
  double *trec_r, *trec_m;
  double *PNpart;
  double *psi_prev;    // [N] psi of the previous FP32 evaluation (history mode)
  double *dpsi_r;      // [N] (psi - psi_prev) * coef, row order (history mode)
  double *psit_r;      // [N] psi + eps ln nu, row order (value bound), or null
  int64_t N;
  int hist;
  double o[3];
  double coef;         // inv_eps * log-base scale
  double eps;
};

struct P1Args {
  const double *qx;      // [nQ*DIM] Morton order, shifted
  const double *qx2;     // [nQ]
  const double *cc;      // [n_cells*DIM] raw centers, Morton order
  const double *blkbox;  // [n_blocks*2*DIM]
  const double *trec;    // [N*kRec] row order
  const int32_t *tstart;
  const double *Mpart;   // [K1*nQ]  maxima of PASS 0, read by PASS 1
  const double *Mglob;   // [nQ]     their maximum over the segments (PASS 1; null: Mpart)
  const double *dpsi;    // [N]      FP32 history: exponent change per target, row order
  const double *dcx, *dcn;  // its maximum / minimum per chunk of 64 records (null: scan dpsi)
  double *Mout;          // [K1*nQ]  PASS 0: segment maxima; PASS 2: reference value used
  double *Sout;          // [K1*nQ]  PASS 1 / 2: segment sums
  double *Tout;          // [K1*nQ]  PASS 2: segment maxima of the exponent (read first:
                         //          the previous evaluation's, for the history seed)
  Grid tg;
  int64_t n_groups, nQ, N;
  int nq, gpc, K1, dense, hist;
  int hist_same;          // the history's radius is this one (else a smaller one)
  double smax, Rq, plain_coef, coef;
  double s_in;            // squared radius of the untested inner runs (0: none)
  double seed_bias;       // added to every seeded reference value (0; a test hook that
                          // exercises the re-sweep of an item whose seed is too high)
  // value bound of the sums (VBound, sdot_device.hpp): off when prune == 0
  const double *qbox;     // [n_blocks*2*DIM] the q-points' box per block (raw coordinates)
  const double *pbin;     // [n_tbins] largest psi + eps ln nu per target bin
  const double *prow;     // [n_trows] ... per bin row
  double o[3];            // shift origin (raw = shifted + o)
  double theta;           // 2^-bits in the sweep's units: bits (FP32) or bits ln 2 (FP64)
  double inv_coef;        // 1 / coef
  double vshift;          // eps ln nu_min
  int prune;
};

struct QFArgs {
  const double *qx, *qw, *qx2, *Mpart, *Spart, *Tpart;
  const int64_t *cell_r2m;
  double *qrec_r, *Fpart, *Cpart, *Upart;
  double *qchunk;  // [nQ*(DIM+2)] the gradient sweep's copy of x', M, w/S (GArgs::qchunk)
  int64_t nQ;
  int nq, K1;
  int G;           // its chunk size (q-points per chunk; divides nq)
  double eps, plain_coef, ln_scale;
};

struct GArgs {
  const double *trec_m;  // [N*kRec] Morton order
  const double *twbox;   // [n_twaves*2*DIM]
  const double *cc_r;    // [n_cells*4] raw centers, row order
  const int32_t *cstart;
  const double *qrec_r;  // [nQ*kRec]
  // x', M and w/S of the q-points in chunks of G (row cell order), each chunk
  // [x'_0 (G values) .. x'_DIM-1, M, w/S]: a chunk's (DIM+2) G doubles are
  // contiguous, so a few wide scalar loads fetch it (the q records take two
  // loads per q-point)
  const double *qchunk;
  double *Gpart;         // [K2*N]
  double *Bpart;         // [K2*DIM*N] first moments sum_q (w/S) e x' (barycenters), or null
  Grid cg;
  int64_t N, n_cells;
  int nq, K2, dense;
  double smax, Rq;
  double s_in;            // squared radius of the untested inner runs (0: none)
  // value bound (VBound, sdot_device.hpp): off when prune == 0
  const double *vcell;    // [n_cells] smallest V_q per cell, row order
  const double *vbin;     // [n_cbins] ... per cell bin
  const double *vrow;     // [n_crows] ... per cell-bin row
  double rq;              // largest distance of a cell's q-points from its center
  double inv_coef, eps;
  int prune;
};

struct GFArgs {
  const double *Gpart, *trec_m;
  const int32_t *tm2o;
  double *grad;
  int64_t N;
  int K2;
};

struct FinArgs {
  const double *Fpart, *Cpart, *Upart, *PNpart;
  int64_t nF, nP;
  double *out, *ucount;
};

struct Profile {
  bool on = false;
  hipEvent_t ev[6] = {};
  sdot_profile acc{};
};

// Shape of a target set's layout, from one host pass over the points
// (sdot_api.cpp): bounding box, coordinate scale, bin grid (about 8 points per
// bin) and the number of 64-target waves of the gradient pass.
struct TargetShape {
  Grid tg{};
  int64_t n = 0, n_twaves = 0, n_bins = 0;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, scale = 0.0;
};
void target_shape(int dim, int64_t n, const double *y, TargetShape &S);

// The same layout built on the host (sdot_api.cpp), for the softmax transfer's
// coarse level, whose records are written on the host.
struct TargetLayout {
  TargetShape shape;
  std::vector<int32_t> m2o, pos_m, r2o, pos_r, tstart;
};
void build_target_layout(int dim, int64_t n, const double *y, TargetLayout &L);

// The target layout built on the device (sdot_layout.hip) from the points in
// caller order already in HBM: stable radix sorts of the row-major bin index
// (q-major record order) and of the Morton key (gradient waves), the bin
// offsets and the wave boxes.  Outputs (N = S.n, dim-strided boxes):
//   pos_r [N] caller -> row position,   r2o [N] row position -> caller (or null),
//   pos_m [N] caller -> Morton position, m2o [N] Morton position -> caller,
//   tstart [n_bins + 1],  twbox [n_twaves * 2 * dim] (min then max per wave),
//   ty_m [N * 4] points in Morton order, padded to 4 (or null).
// scratch: layout_scratch_bytes(S) bytes of device memory.
size_t layout_scratch_bytes(const TargetShape &S);
hipError_t device_target_layout(int dim, const TargetShape &S, const double *d_y, int32_t *pos_r,
                                int32_t *r2o, int32_t *pos_m, int32_t *m2o, int32_t *tstart,
                                double *twbox, double *ty_m, void *scratch, size_t scratch_bytes,
                                hipStream_t s);

// Largest double s with sqrt(s) <= R (candidate test s <= smax).
double smax_for(double R);
// Squared radius of the untested inner runs for the bound smax (0: none).
double inner_s(double smax);
// Sets the thread's sdot_last_error message; returns code.
int set_error(int code, const std::string &msg);

// Read-only view of a context's source arrays (serving queues borrow them).
struct SourceView {
  int device = 0, dim = 0, nq = 0, G = 1, gpc = 1;
  int64_t n_cells = 0, nQ = 0, n_groups = 0, n_blocks = 0;
  Grid cg{};
  double origin[3] = {0, 0, 0}, lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, coord_scale = 1.0;
  const double *qx_m = nullptr, *qw_m = nullptr, *qx2_m = nullptr, *cc_m = nullptr;
  const double *blkbox = nullptr, *cc_r = nullptr;
  const int32_t *cstart = nullptr;
  const int64_t *cell_r2m = nullptr;
  // the 4-point grouping of the value-bounded sweeps, and the bound's cell data
  int Gs = 1, gpc_s = 1;
  int64_t n_groups_s = 0, n_blocks_s = 0, n_cbins = 0;
  const double *blkbox_s = nullptr, *qbox_s = nullptr;
  double rq = 0.0;
};
// Value bound of the pair sums: FP32_LSE evaluations are bounded at this many
// bits by default where the bound pays (value_bound_default below): terms below
// 2^-50 of their q-point's largest, far below the 2^-24 rounding of the FP32
// exponentials.
constexpr double kPruneBitsF32 = 50.0;
// Whether the FP32_LSE default bounds an evaluation of N targets against nQ
// quadrature points at epsilon and truncation radius R, with source-box
// diagonal diag: the radius sqrt(2 bits ln2 eps) beyond which a target's terms
// fall below 2^-bits (at equal potentials) is at most a fifth of the diagonal
// and half of R, and there are at least 2^32 pairs.  Where the radius is close
// to R or to the source's extent the bound skips little and its sweeps (4-point
// lanes, fewer waves per SIMD) cost more: the 500,000-target ball workload
// (R = 0.205, radius 0.186) solved in 19.9 s bounded against 17.9 s, and served
// trace requests at eps 3e-3 (radius 0.46 in the unit cube) took longer.
inline bool value_bound_default(double eps, double diag, double R, int64_t N, int64_t nQ) {
  const double r0 = std::sqrt(2.0 * kPruneBitsF32 * 0.69314718055994530942 * eps);
  return r0 <= 0.2 * diag && !(r0 > 0.5 * R) && (double)N * (double)nQ >= 4294967296.0;
}
int ctx_source_view(sdot_ctx *c, SourceView &v);
// Count of serving queues attached to the context (its source stays fixed meanwhile).
void ctx_attach(sdot_ctx *c, int delta);

// Rounding noise of the assembled functional relative to |F|, for the
// noise-aware line search: FP64 evaluations scatter by ~2e-16 |F| and FP32_LSE
// evaluations (FP32 exponentials) by ~4e-10 |F| between neighbouring
// potentials; the constants keep a factor 5-10 above that.
constexpr double kNoiseRelF64 = 2e-15;
constexpr double kNoiseRelF32 = 2e-9;
// FP32_LSE solves switch to FP64 evaluations once their stopping value is
// within this factor of the tolerance (the serving queue's default).
constexpr double kPolishFactor = 2.0;

// Softmax potential transfer from a coarse to a fine target level
// (SoftmaxRefinement::compute_refinement, SoftmaxRefinement.cc:61-280), FP64.
// Entries are (coarse parent, fine child) pairs of the level's child lists,
// sorted by the Morton position of the parent.
struct RefArgs {
  const double *qxr;      // [nQ*DIM] raw quadrature points, Morton cell order
  const double *qw;       // [nQ] rho*JxW, Morton cell order
  const double *cc;       // [n_cells*DIM] raw centers, Morton order
  const double *cc_r;     // [n_cells*4] raw centers, row (bin) order
  const int32_t *cstart;  // row-order cell offsets per cell bin
  const int64_t *cell_r2m;
  const double *crec;     // [Nc*kRec] coarse points, row (bin) order: y[DIM] at 0, psi at 3, nu at 4
  const int32_t *tstart;  // coarse bin offsets
  const double *erec;     // [E*kRec] entries: parent y[DIM] at 0, child y[DIM] at 3
  const int64_t *foff;    // [n_fine+1] fine point -> its entries (CSR, caller order)
  const int64_t *fent;    // [E]
  double *Mpart, *Spart;  // [K1*nQ] Morton cell order
  double *rq;             // [nQ*4] row cell order: x[DIM] at 0, w e^{-M} / S at 3 (0: skipped q)
  double *Epart;          // [K2*E]
  double *psi_fine;       // [n_fine]
  Grid tg, cg;
  int64_t nQ, n_cells, Nc, E, n_fine;
  int nq, K1, K2, lse;
  double smax;   // cell rule:  |y_c - center| <= R   <=>  s <= smax
  double s2max;  // per-q rule: pow(|x_q - y|, 2) <= R*R  <=>  s <= s2max
  double Rq, lam, lam_inv;
};
// Largest double s with pow(sqrt(s), 2) <= R*R (the per-q test of
// SoftmaxRefinement.cc:152,162,173,192).
double s2max_for(double R);
hipError_t launch_refinement(int dim, const RefArgs &a, hipStream_t s);

// Point queries over a context's target measure (sdot_points.hip): conditional
// densities (SotSolver::get_potential_conditioned_density, SotSolver.cc:973-1046)
// and transport maps (OptimalTransportPlan.cc:99-189, 218-310), FP64.
struct PtArgs {
  const double *px;       // [n*DIM] query points, Morton order
  const double *praw;     // [n*DIM] query points, caller order
  const int64_t *pm2o;    // [n] Morton position -> caller index
  const double *rec;      // [N*kRec] targets, row (bin) order (written by k_pt_rec)
  double *rec_w;
  const int32_t *tstart;  // target bin offsets
  const int32_t *tpos_r;  // caller index -> row position
  const double *ty, *nu, *psi;  // caller order
  const double *logterm;  // [N] eps log nu_j (-inf where nu_j <= 0), modal map only
  Grid tg;
  int64_t n, N;
  int K;
  int divide;     // exponent (psi - sq/2) / eps instead of * (1/eps)
  int moments;    // sum the first moments (barycentric map)
  int truncated;  // points without a candidate take the nearest target
  double smax;    // candidate test s <= smax
  double Rq;      // enumeration radius
  double eps, eps_inv;
  double rho0, rho_all;  // nearest-target scans
  double *Mpart, *Spart, *Bpart;  // [K*n], [K*n], [K*DIM*n], Morton order
  int64_t *Cpart, *Ipart;         // [K*n] candidate counts, modal indices
  double *Mo, *So;                // [n] caller order (conditional densities)
  int32_t *need;                  // [n] Morton order: nearest-target fallback
  double *mapped;                 // [n*DIM] caller order
  int64_t n_idx;
  const int64_t *idx;             // [n_idx] requested target indices
  const double *rho;              // [n] density at the points, caller order
  double *cond;                   // [n_idx*n]
};
// op 0: conditional densities, 1: barycentric map, 2: modal map.
hipError_t launch_points(int dim, const PtArgs &a, int op, hipStream_t s);

// The evaluation that allocates nothing (sdot_eval_device_noalloc) carves its
// per-evaluation buffers from one file-scope device array of this many bytes
// per GPU (sdot_kernels.hip), which each GPU holds until the process ends.
constexpr size_t kNoAllocBytes = size_t(128) << 20;
// Device address of that array on `device`, which must be the current device;
// looked up with hipGetSymbolAddress on the first call per device and kept.
hipError_t noalloc_scratch(int device, double **ptr);

// Kernel launchers (sdot_kernels.hip).
hipError_t launch_eval(const Plan &p, Dev &d, const double *d_psi, double *d_grad,
                       double *d_FC, hipStream_t s, Profile *prof);
// The same evaluation with the first moments of the transport plan per target:
// d_bary [N*dim] (caller order) = sum_q (w_q / S_q) e_qj x_q, d_Bpart scratch [K2*dim*N].
hipError_t launch_eval_bary(const Plan &p, Dev &d, const double *d_psi, double *d_grad,
                            double *d_FC, double *d_bary, double *d_Bpart, hipStream_t s);
hipError_t launch_candidate_stats(const Plan &p, Dev &d, int64_t *d_count,
                                  uint64_t *d_hash, hipStream_t s);
hipError_t launch_covering(const Plan &p, Dev &d, double *d_cellmin, hipStream_t s);
hipError_t launch_hessian_chunk(const Plan &p, Dev &d, int64_t q_begin, int64_t rows,
                                double *d_P, double *d_A, int64_t Npad, hipStream_t s);
hipError_t launch_hessian_final(const Plan &p, const double *d_A, int64_t Npad, const double *d_g,
                                const double *d_nu, double *d_H, hipStream_t s);

// L-BFGS vector arithmetic on device vectors of length n (sdot_lbfgs.hip).
// Reductions leave kLbParts partial values per quantity in `part` (a fixed
// grid, so the host's in-order sum of them does not depend on the launch).
constexpr int kLbParts = 512;
constexpr int kLbMaxHist = 16;
// part[0, kLbParts) = partial sums of a . b
hipError_t launch_lb_dot(const double *a, const double *b, int64_t n, double *part, hipStream_t s);
// x = x0 + a p
hipError_t launch_lb_lin(double *x, const double *x0, double a, const double *p, int64_t n,
                         hipStream_t s);
// s = x - s, y = g - y
hipError_t launch_lb_diff(double *sv, const double *x, double *y, const double *g, int64_t n,
                          hipStream_t s);
// kLbStats blocks of kLbParts partial values in part, for vector a (and nu):
// sum |a|, sum a^2, max (|a| - t nu) (nu == nullptr: -inf), max a, min a
enum { kLbL1 = 0, kLbL2, kLbCw, kLbMax, kLbMin, kLbStats };
hipError_t launch_lb_stats(const double *a, const double *nu, double t, int64_t n, double *part,
                           hipStream_t s);
// Two-loop recursion: p holds g on entry and -H g on return, H the
// inverse-Hessian approximation of the pairs (S[k], Y[k]) for k in
// slots[0..count) (newest first, with their rho) over the initial diag(h0)
// (h0 == nullptr: identity).  part: kLbParts partial sums; coef: count values.
struct LbVecs {
  const double *S[kLbMaxHist], *Y[kLbMaxHist];
  int64_t n;
  double *p;
  const double *h0;
  double *part, *coef;
};
hipError_t launch_lb_direction(const LbVecs &L, const int *slots, const double *rho, int count,
                               hipStream_t s);

}  // namespace sdot
