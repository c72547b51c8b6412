// sdot_kernels.hip -- gfx950 kernels of the regularized semi-discrete OT dual
// evaluation (the work of SotSolver::evaluate_functional / local_assemble,
// SotSolver.cc:217-377 and 769-869).
//
// One evaluation is a handful of launches on one stream:
//   k_prep     per target: exponent coefficients a_j = (psi_j - |y'_j|^2/2)/eps and
//              y'_j/eps (scaled by log2(e) in the FP32 mode), written in the two
//              target orders the passes stream; block partials of psi.nu and of
//              the change of psi since the previous evaluation.
//   k_pass1    q-major sweep over each cell's candidates.  Lane = group of G
//              quadrature points of one cell; wave = 64 Morton-consecutive groups;
//              grid.y splits the candidate bin rows into K1 segments.
//                PASS 0: max exponent (FP64 LSE),
//                PASS 1: sum_j nu_j exp(t_qj - M_q) with the global max (FP64 modes),
//                PASS 2: FP32 sweep: S = sum nu exp2(t - M) against an integer
//                        upper bound M of the item's exponents (its exact
//                        maximum, found by a max loop first, or the previous
//                        evaluation's maximum raised by the largest exponent
//                        change among its targets), and the running maximum of t.
//   k_qfin     per q: merge segments in fixed order, F_q, w_q/S_q, C_q, the q record
//              the gradient pass streams; block partials of F and C.
//   k_grad     target-major: lane = target, loops over the cells near its wave and
//              their q records; per-segment partial sums, no atomics.
//   k_gfin / k_final  fixed-order segment merge, linear term, scalar totals.
// Every reduction has a fixed order, so results are bitwise reproducible run to
// run and independent of scheduling.
//
// Candidate rule (find_nearest_target_points, SotSolver.cc:553-570, with
// euclidean_distance, Distance.h:23-27): target j belongs to cell K iff
// sqrt(((y0-c0)^2 + (y1-c1)^2) + (y2-c2)^2) <= R, evaluated in FP64 without
// FMA contraction.  Because sqrt is correctly rounded and monotone this equals
// s <= smax with smax the largest double whose sqrt is <= R (host-computed).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "sdot_device.hpp"

namespace sdot {

// ---------------------------------------------------------------------------
// scratch of the evaluation that allocates nothing (sdot_eval_device_noalloc)
// ---------------------------------------------------------------------------
// Each GPU holds its own copy at its own address.  In host code the name is a
// host-side stand-in, never a device address: host code reaches the array
// only through the address noalloc_scratch looks up.
static __device__ __align__(256) double scratch[kNoAllocBytes / sizeof(double)];

hipError_t noalloc_scratch(int device, double **ptr) {
  static std::mutex mu;
  static std::vector<double *> addr;  // per device; null until looked up
  if (device < 0 || !ptr) return hipErrorInvalidValue;
  std::lock_guard<std::mutex> lock(mu);
  if ((size_t)device >= addr.size()) addr.resize((size_t)device + 1, nullptr);
  if (!addr[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    addr[device] = static_cast<double *>(p);
  }
  *ptr = addr[device];
  return hipSuccess;
}

// ---------------------------------------------------------------------------
// single-solve entry points (grid = the solve's own blocks)
// ---------------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(256) void k_prep(PrepArgs a) {
  __shared__ double sh[256];
  prep_body<DIM>(a, blockIdx.x, sh);
}

// value bound: per-bin and per-row reductions, per-cell bounds
// one wave per bin, four bins per workgroup
template <bool MAX>
__global__ __launch_bounds__(256) void k_binred(const double *v, const int32_t *start, int64_t nb,
                                                double *out) {
  bin_reduce_body<MAX>(v, start, nb, out, (int64_t)blockIdx.x * 4 + threadIdx.x / kWave);
}

template <bool MAX>
__global__ __launch_bounds__(256) void k_rowred(const double *bins, int64_t nrows, int nx,
                                                double *out) {
  row_reduce_body<MAX>(bins, nrows, nx, out, (int64_t)blockIdx.x * 256 + threadIdx.x);
}

template <int DIM>
__global__ __launch_bounds__(256) void k_cellv(const double *qrec_r, int64_t n_cells, int nq,
                                               double theta, double inv_coef, int log2_units,
                                               double *vcell) {
  cellv_body<DIM>(qrec_r, n_cells, nq, theta, inv_coef, log2_units != 0, vcell,
                  (int64_t)blockIdx.x * 256 + threadIdx.x);
}

static unsigned nblk256(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

__global__ __launch_bounds__(256) void k_chunkext(const double *dpsi_r, int64_t N, double *dcx,
                                                  double *dcn) {
  chunkext_body(dpsi_r, N, dcx, dcn, (int64_t)blockIdx.x * (256 / kWave) + threadIdx.x / kWave);
}

// At most 4 points per lane (truncated sweeps): 5 waves per SIMD, which the
// unbounded sweeps reach without spills (their records stream two at a time).
template <int DIM, int G, int MODE, int PASS, bool VB>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu((G <= 4 && !VB) ? 5 : 1)))
void k_pass1(P1Args a) {
  pass1_body<DIM, G, MODE, PASS, VB>(a, blockIdx.x, blockIdx.y);
}

// Maximum of the K1 segment maxima per quadrature point (FP64 LSE), read once
// by each item of PASS 1 instead of every segment's.
__global__ __launch_bounds__(256) void k_mmerge(const double *Mpart, int64_t nQ, int K1,
                                                double *Mglob) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= nQ) return;
  Mglob[q] = fold_partials(Mpart + q, nQ, K1, -INFINITY, FoldMax{});
}

template <int DIM, int MODE>
__global__ __launch_bounds__(256) void k_qfin(QFArgs a) {
  __shared__ double sh[256];
  qfin_body<DIM, MODE>(a, blockIdx.x, sh);
}

// One-dimensional grid, segment index fastest.  Workgroups are dispatched
// round-robin over the 8 XCDs, so with K2 a multiple of 8 XCD x runs exactly
// the segments s = x (mod 8): its L2 holds one eighth of the q records instead
// of streaming all of them.
template <int DIM, int MODE, int G, bool VB>
__global__ __launch_bounds__(64) void k_grad(GArgs a) {
  const int64_t blk = blockIdx.x / a.K2;
  grad_body<DIM, MODE, G, false, VB>(a, blk, (int)(blockIdx.x - blk * a.K2));
}

// The gradient sweep with the first moments of the plan (barycenters).
template <int DIM, int MODE, int G, bool VB>
__global__ __launch_bounds__(64) void k_grad_bary(GArgs a) {
  const int64_t blk = blockIdx.x / a.K2;
  grad_body<DIM, MODE, G, true, VB>(a, blk, (int)(blockIdx.x - blk * a.K2));
}

struct Origin {
  double v[3];
};

template <int DIM>
__global__ __launch_bounds__(256) void k_bfin(const double *Gpart, const double *Bpart,
                                              const int32_t *tm2o, Origin o, int64_t N, int K2,
                                              double *bary) {
  bfin_body<DIM>(Gpart, Bpart, tm2o, o.v, N, K2, bary, blockIdx.x);
}

template <int DIM>
__global__ __launch_bounds__(256) void k_gfin(GFArgs a) {
  gfin_body<DIM>(a, blockIdx.x);
}

__global__ __launch_bounds__(256) void k_final(FinArgs a) {
  __shared__ double sh[256];
  final_body(a, sh);
}

// ---------------------------------------------------------------------------
// audit: candidate-set count and hash per cell
// ---------------------------------------------------------------------------
struct CSArgs {
  const double *cc, *blkbox, *trec;
  const int32_t *tstart, *tr2o;
  const int64_t *cell_m2o;
  int64_t *count;
  uint64_t *hash;
  Grid tg;
  int64_t n_groups, N;
  int gpc, dense;
  double smax, Rq;
};

template <int DIM>
__global__ __launch_bounds__(64) void k_cstats(CSArgs a) {
  const int64_t grp = (int64_t)blockIdx.x * kWave + threadIdx.x;
  const bool active = grp < a.n_groups;
  const int64_t gg = active ? grp : 0;
  const int64_t cell = gg / a.gpc;
  const bool owner = active && (gg % a.gpc) == 0;
  double c[DIM], lo[DIM], hi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    c[d] = a.cc[cell * DIM + d];
    lo[d] = a.blkbox[(int64_t)blockIdx.x * 2 * DIM + d];
    hi[d] = a.blkbox[(int64_t)blockIdx.x * 2 * DIM + DIM + d];
  }
  int64_t cnt = 0;
  uint64_t h = 0;
  // the audit applies the exact rule to every enumerated target (no inner runs)
  auto body = [&](int j0, int j1, auto) {
    for (int j = j0; j < j1; ++j) {
      const double *r = a.trec + (int64_t)j * kRec;
      double yr[DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) yr[d] = r[1 + DIM + d];
      if (!a.dense && !in_set<DIM>(yr, c, a.smax)) continue;
      uint64_t v = (uint64_t)a.tr2o[j] + 1u;
      v *= 0x9E3779B97F4A7C15ull;
      v ^= v >> 29;
      h += v;
      ++cnt;
    }
  };
  if (a.dense)
    body(0, (int)a.N, Inside{});
  else
    for_rows<DIM>(a.tg, a.tstart, lo, hi, a.Rq, 0.0, 0, 1, body);
  if (owner) {
    const int64_t oc = a.cell_m2o[cell];
    a.count[oc] = cnt;
    a.hash[oc] = h;
  }
}

// Covering radius: per cell (Morton order) the min over all targets of the
// squared center distance, same arithmetic as the candidate rule.
template <int DIM>
__global__ __launch_bounds__(256) void k_cover(const double *cc, const double *ty_m,
                                               int64_t n_cells, int64_t N, double *cellmin) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const bool active = c < n_cells;
  double cv[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) cv[d] = cc[(active ? c : 0) * DIM + d];
  double best = INFINITY;
  for (int64_t j = 0; j < N; ++j) {
#pragma clang fp contract(off)
    double s = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      const double t = ty_m[j * 4 + d] - cv[d];
      s = s + t * t;
    }
    best = fmin(best, s);
  }
  if (active) cellmin[c] = best;
}

// ---------------------------------------------------------------------------
// Dense Hessian  H = (1/eps) [ diag(g + nu) - P^T P ],
//   P[q][j] = sqrt(w_q) p_qj,  p_qj = nu_j exp(t_qj - M_q) / S_q  (0 off C(cell(q)))
// built in chunks of quadrature rows; P^T P by FP64 MFMA (16x16x4), upper
// block triangle only, accumulated chunk after chunk in a fixed order.
// ---------------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(256) void k_pbuild(const double *trec_m, const int32_t *pos_m,
                                                const double *qrec_r, const double *cc_r,
                                                int64_t N, int64_t Npad, int nq, int64_t q_begin,
                                                int64_t rows, int64_t nQ, int dense, double smax,
                                                double *P) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= Npad) return;
  double y[DIM], yl[DIM], ta = 0.0, nu = 0.0;
  if (j < N) {
    const double *r = trec_m + (int64_t)pos_m[j] * kRec;
    ta = r[0];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      yl[d] = r[1 + d];
      y[d] = r[1 + DIM + d];
    }
    nu = r[1 + 2 * DIM];
  }
  for (int64_t row = blockIdx.y; row < rows; row += gridDim.y) {
    const int64_t qr = q_begin + row;
    double v = 0.0;
    if (j < N && qr < nQ) {
      const double *qp = qrec_r + qr * kRec;
      const double scale = qp[DIM + 1];
      if (scale > 0.0) {
        bool in = true;
        if (!dense) {
          double c[DIM];
#pragma unroll
          for (int d = 0; d < DIM; ++d) c[d] = cc_r[(qr / nq) * 4 + d];
          in = in_set<DIM>(y, c, smax);
        }
        if (in) {
          double t = ta;
#pragma unroll
          for (int d = 0; d < DIM; ++d) t = fma(qp[d], yl[d], t);
          v = nu * exp(t - qp[DIM]) * sqrt(scale / qp[DIM + 2]);
        }
      }
    }
    P[row * Npad + j] = v;
  }
}

typedef double sdot_d4 __attribute__((ext_vector_type(4)));
typedef double sdot_d2 __attribute__((ext_vector_type(2)));

// A[j][k] += sum_q P[q][j] P[q][k] on 128x128 blocks with bk >= bj (kSyT
// columns per side).  Eight waves per block (two waves per SIMD at 118 VGPRs
// and 64 KB of LDS per block: two blocks per CU), each wave a 64x32 piece as
// 4x2 MFMA 16x16x4 tiles.  The P rows stream through LDS in stages of kSyKS
// rows, double-buffered: stage s+1 is loaded into registers while stage s
// feeds the MFMAs.  LDS holds a stage MFMA-ready: for every K-step of 4 rows
// and every 16-column tile, the 64 operands in lane order (lane l: row l/16,
// column l%16), so each operand fetch of a wave is 64 consecutive doubles.
// Measured on MI355X (P^T P, 16,384 / 32,768 columns): 53 / 58 TFLOP/s,
// against 35 / 38 for 64x64 blocks of four waves, with or without LDS.
constexpr int kSyT = 128;
constexpr int kSyKS = 16;
constexpr int kSyThreads = 512;
constexpr int kSyStage = kSyKS * kSyT;                     // doubles per side and stage
constexpr int kSyPairs = kSyStage / 2 / kSyThreads;        // double2 loads per thread and side

__device__ __forceinline__ void syrk_fetch(const double *__restrict__ P, int64_t Npad,
                                           int64_t q0, int64_t c0, int tid,
                                           sdot_d2 (&r)[kSyPairs]) {
#pragma unroll
  for (int i = 0; i < kSyPairs; ++i) {
    const int o = 2 * (i * kSyThreads + tid);    // LDS offset of the pair in the stage
    const int grp = o >> 6, slot = o & 63;       // (K-step, tile), lane slot
    const int row = (grp >> 3) * 4 + (slot >> 4), col = (grp & 7) * 16 + (slot & 15);
    r[i] = *reinterpret_cast<const sdot_d2 *>(P + (q0 + row) * Npad + c0 + col);
  }
}

__device__ __forceinline__ void syrk_store(double *L, int tid, const sdot_d2 (&r)[kSyPairs]) {
#pragma unroll
  for (int i = 0; i < kSyPairs; ++i) *reinterpret_cast<sdot_d2 *>(L + 2 * (i * kSyThreads + tid)) = r[i];
}

// rows: a multiple of kSyKS; Npad: a multiple of kSyT.
__global__ __launch_bounds__(kSyThreads) void k_syrk(const double *__restrict__ P, int64_t rows,
                                                     int64_t Npad, double *__restrict__ A) {
  constexpr int TJ = 4, TK = 2;               // MFMA tiles per wave along j and k
  constexpr int WJN = kSyT / (16 * TJ);       // waves along j (2); along k: 4
  const int bj = blockIdx.y, bk = blockIdx.x;
  if (bk < bj) return;
  const bool diag = bj == bk;
  __shared__ __align__(16) double lds[2][2][kSyStage];
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int wj = w % WJN, wk = w / WJN;
  const int64_t cj = (int64_t)bj * kSyT, ck = (int64_t)bk * kSyT;
  sdot_d4 acc[TJ][TK];
#pragma unroll
  for (int t = 0; t < TJ; ++t)
#pragma unroll
    for (int u = 0; u < TK; ++u) acc[t][u] = sdot_d4{0.0, 0.0, 0.0, 0.0};
  sdot_d2 rj[kSyPairs], rk[kSyPairs];
  syrk_fetch(P, Npad, 0, cj, tid, rj);
  if (!diag) syrk_fetch(P, Npad, 0, ck, tid, rk);
  syrk_store(lds[0][0], tid, rj);
  if (!diag) syrk_store(lds[0][1], tid, rk);
  __syncthreads();
  const int64_t stages = rows / kSyKS;
  for (int64_t st = 0; st < stages; ++st) {
    const int buf = (int)(st & 1);
    const bool more = st + 1 < stages;
    if (more) {
      syrk_fetch(P, Npad, (st + 1) * kSyKS, cj, tid, rj);
      if (!diag) syrk_fetch(P, Npad, (st + 1) * kSyKS, ck, tid, rk);
    }
    const double *Lj = lds[buf][0];
    const double *Lk = diag ? lds[buf][0] : lds[buf][1];
#pragma unroll
    for (int ks = 0; ks < kSyKS / 4; ++ks) {
      double a[TJ], b[TK];
#pragma unroll
      for (int t = 0; t < TJ; ++t) a[t] = Lj[(ks * 8 + wj * TJ + t) * 64 + l];
#pragma unroll
      for (int u = 0; u < TK; ++u) b[u] = Lk[(ks * 8 + wk * TK + u) * 64 + l];
#pragma unroll
      for (int t = 0; t < TJ; ++t)
#pragma unroll
        for (int u = 0; u < TK; ++u)
          acc[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b[u], acc[t][u], 0, 0, 0);
    }
    if (more) {
      // the other buffer was last read in stage st-1, before that stage's barrier
      syrk_store(lds[buf ^ 1][0], tid, rj);
      if (!diag) syrk_store(lds[buf ^ 1][1], tid, rk);
    }
    __syncthreads();
  }
  const int lr = l >> 4, lc = l & 15;
#pragma unroll
  for (int t = 0; t < TJ; ++t)
#pragma unroll
    for (int u = 0; u < TK; ++u)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        // f64 16x16x4 output: register i of lane l holds row 4*i + l/16, column l%16
        const int64_t jj = cj + wj * TJ * 16 + t * 16 + i * 4 + lr;
        const int64_t kk = ck + wk * TK * 16 + u * 16 + lc;
        A[jj * Npad + kk] += acc[t][u][i];
      }
}

__global__ __launch_bounds__(256) void k_hfin(const double *A, int64_t Npad, const double *g,
                                              const double *nu, int64_t N, double inv_eps,
                                              double *H) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= N) return;
  for (int64_t j = blockIdx.y; j < N; j += gridDim.y) {
    const int64_t lo = j < k ? j : k, hi = j < k ? k : j;
    const double diag = (j == k) ? g[j] + nu[j] : 0.0;
    H[j * N + k] = (diag - A[lo * Npad + hi]) * inv_eps;
  }
}

hipError_t launch_hessian_chunk(const Plan &p, Dev &d, int64_t q_begin, int64_t rows, double *d_P,
                                double *d_A, int64_t Npad, hipStream_t s) {
  const dim3 gb((unsigned)((Npad + 255) / 256), (unsigned)std::min<int64_t>(rows, 4096));
  if (p.dim == 2)
    hipLaunchKernelGGL(k_pbuild<2>, gb, dim3(256), 0, s, d.trec_m, d.tpos_m, d.qrec_r, d.cc_r,
                       p.N, Npad, p.nq, q_begin, rows, p.nQ, p.dense, p.smax, d_P);
  else
    hipLaunchKernelGGL(k_pbuild<3>, gb, dim3(256), 0, s, d.trec_m, d.tpos_m, d.qrec_r, d.cc_r,
                       p.N, Npad, p.nq, q_begin, rows, p.nQ, p.dense, p.smax, d_P);
  if (Npad % kSyT != 0 || rows % kSyKS != 0) return hipErrorInvalidValue;
  const unsigned nb = (unsigned)(Npad / kSyT);
  hipLaunchKernelGGL(k_syrk, dim3(nb, nb), dim3(kSyThreads), 0, s, d_P, rows, Npad, d_A);
  return hipGetLastError();
}

hipError_t launch_hessian_final(const Plan &p, const double *d_A, int64_t Npad, const double *d_g,
                                const double *d_nu, double *d_H, hipStream_t s) {
  const dim3 g((unsigned)((p.N + 255) / 256), (unsigned)std::min<int64_t>(p.N, 4096));
  hipLaunchKernelGGL(k_hfin, g, dim3(256), 0, s, d_A, Npad, d_g, d_nu, p.N, 1.0 / p.eps, d_H);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

template <int DIM, int G, int MODE, int PASS>
static void launch_pass1_t(const P1Args &a, int64_t nb, int K1, hipStream_t s) {
  const dim3 grid((unsigned)nb, (unsigned)K1);
  // the value-bounded sweeps run the small grouping (at most 4 points per lane)
  if (G <= 4 && a.prune)
    hipLaunchKernelGGL((k_pass1<DIM, (G <= 4 ? G : 1), MODE, PASS, true>), grid, dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL((k_pass1<DIM, G, MODE, PASS, false>), grid, dim3(64), 0, s, a);
}

template <int DIM, int MODE, int PASS>
static hipError_t launch_pass1_g(int G, const P1Args &a, int64_t nb, int K1, hipStream_t s) {
  if (a.prune && G > 4) return hipErrorInvalidValue;
  switch (G) {
    case 1: launch_pass1_t<DIM, 1, MODE, PASS>(a, nb, K1, s); break;
    case 2: launch_pass1_t<DIM, 2, MODE, PASS>(a, nb, K1, s); break;
    case 3: launch_pass1_t<DIM, 3, MODE, PASS>(a, nb, K1, s); break;
    case 4: launch_pass1_t<DIM, 4, MODE, PASS>(a, nb, K1, s); break;
    case 5: launch_pass1_t<DIM, 5, MODE, PASS>(a, nb, K1, s); break;
    case 6: launch_pass1_t<DIM, 6, MODE, PASS>(a, nb, K1, s); break;
    case 7: launch_pass1_t<DIM, 7, MODE, PASS>(a, nb, K1, s); break;
    case 8: launch_pass1_t<DIM, 8, MODE, PASS>(a, nb, K1, s); break;
    case 9: launch_pass1_t<DIM, 9, MODE, PASS>(a, nb, K1, s); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int DIM, int MODE, bool BARY>
static hipError_t launch_grad_g(int G, const GArgs &g, int64_t ntw, int K2, hipStream_t s) {
  const dim3 gg((unsigned)(ntw * K2));
  if (g.prune && G > 4) return hipErrorInvalidValue;
#define SDOT_GRAD(GG)                                                                          \
  do {                                                                                         \
    constexpr int GV = GG <= 4 ? GG : 1;                                                       \
    if constexpr (BARY) {                                                                      \
      if (g.prune)                                                                             \
        hipLaunchKernelGGL((k_grad_bary<DIM, MODE, GV, true>), gg, dim3(64), 0, s, g);         \
      else                                                                                     \
        hipLaunchKernelGGL((k_grad_bary<DIM, MODE, GG, false>), gg, dim3(64), 0, s, g);        \
    } else {                                                                                   \
      if (g.prune)                                                                             \
        hipLaunchKernelGGL((k_grad<DIM, MODE, GV, true>), gg, dim3(64), 0, s, g);              \
      else                                                                                     \
        hipLaunchKernelGGL((k_grad<DIM, MODE, GG, false>), gg, dim3(64), 0, s, g);             \
    }                                                                                          \
  } while (0)
  switch (G) {
    case 1: SDOT_GRAD(1); break;
    case 2: SDOT_GRAD(2); break;
    case 3: SDOT_GRAD(3); break;
    case 4: SDOT_GRAD(4); break;
    case 5: SDOT_GRAD(5); break;
    case 6: SDOT_GRAD(6); break;
    case 7: SDOT_GRAD(7); break;
    case 8: SDOT_GRAD(8); break;
    default: SDOT_GRAD(9); break;
  }
#undef SDOT_GRAD
  return hipGetLastError();
}

template <int DIM, int MODE>
static hipError_t eval_t(const Plan &p, Dev &d, const double *d_psi, double *d_grad,
                         double *d_FC, hipStream_t s, Profile *prof,
                         double *d_bary = nullptr, double *d_Bpart = nullptr) {
  const double ln_scale = (MODE == kF32Lse) ? 0.69314718055994530942 : 1.0;
  const double coef = (1.0 / p.eps) * ((MODE == kF32Lse) ? 1.44269504088896340736 : 1.0);
  const bool timing = prof && prof->on;
  if (timing) (void)hipEventRecord(prof->ev[0], s);

  PrepArgs pa{};
  pa.y = d.ty_o;
  pa.nu = d.tnu_o;
  pa.psi = d_psi;
  pa.pos_r = d.tpos_r;
  pa.pos_m = d.tpos_m;
  pa.trec_r = d.trec_r;
  pa.trec_m = d.trec_m;
  pa.PNpart = d.PNpart;
  pa.psi_prev = (MODE == kF32Lse) ? d.psi_prev : nullptr;
  pa.dpsi_r = d.dpsi_r;
  pa.N = p.N;
  pa.hist = p.hist;
  for (int k = 0; k < 3; ++k) pa.o[k] = p.origin[k];
  pa.coef = coef;
  pa.eps = p.eps;
  const bool prune = p.prune_bits > 0 && MODE != kF64Plain;
  pa.psit_r = prune ? d.psit_r : nullptr;
  hipLaunchKernelGGL((k_prep<DIM>), dim3((unsigned)p.nPpart), dim3(256), 0, s, pa);
  // value bound in the sweep's units: 2^-bits is bits log2 units (FP32) or bits ln 2 nats
  const double theta = prune ? p.prune_bits * ((MODE == kF32Lse) ? 1.0 : 0.69314718055994530942) : 0.0;
  if (prune) {
    const int64_t ntr = grid_rows(p.tg);
    hipLaunchKernelGGL(k_binred<true>, dim3((unsigned)((p.n_tbins + 3) / 4)), dim3(256), 0, s,
                       d.psit_r, d.tstart, p.n_tbins, d.pbin);
    hipLaunchKernelGGL(k_rowred<true>, dim3(nblk256(ntr)), dim3(256), 0, s, d.pbin, ntr,
                       p.tg.n[0], d.prow);
  }
  const bool chunks = MODE == kF32Lse && p.hist;
  if (chunks) {
    const int64_t nch = (p.N + kChunk - 1) / kChunk;
    hipLaunchKernelGGL(k_chunkext, dim3((unsigned)((nch + 3) / 4)), dim3(256), 0, s, d.dpsi_r, p.N,
                       d.dpsi_cx, d.dpsi_cn);
  }

  P1Args a{};
  a.qx = d.qx_m;
  a.qx2 = d.qx2_m;
  a.cc = d.cc_m;
  a.blkbox = p.small ? d.blkbox_s : d.blkbox;
  a.trec = d.trec_r;
  a.tstart = d.tstart;
  a.Mpart = d.Mpart;
  a.dpsi = d.dpsi_r;
  a.dcx = chunks ? d.dpsi_cx : nullptr;
  a.dcn = chunks ? d.dpsi_cn : nullptr;
  a.tg = p.tg;
  a.n_groups = p.n_groups;
  a.nQ = p.nQ;
  a.N = p.N;
  a.nq = p.nq;
  a.gpc = p.gpc;
  a.K1 = p.K1;
  a.dense = p.dense;
  a.hist = p.hist;
  a.hist_same = p.hist_same;
  a.smax = p.smax;
  a.Rq = p.Rq;
  a.s_in = p.s_in;
  a.seed_bias = p.seed_bias;
  a.plain_coef = 0.5 / p.eps;
  a.coef = coef;
  a.Mout = d.Mpart;
  a.Sout = d.Spart;
  a.Tout = d.Tpart;
  a.prune = prune ? 1 : 0;
  a.qbox = d.qbox_s;
  a.pbin = d.pbin;
  a.prow = d.prow;
  for (int k = 0; k < 3; ++k) a.o[k] = p.origin[k];
  a.theta = theta;
  a.inv_coef = 1.0 / coef;
  a.vshift = prune ? p.eps * std::log(p.nu_min) : 0.0;
  hipError_t e;
  // exact maxima first (FP64 LSE; the FP32 sweep finds its own bounds)
  if (MODE == kF64Lse) {
    e = launch_pass1_g<DIM, MODE, 0>(p.G, a, p.n_blocks, p.K1, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_mmerge, dim3((unsigned)((p.nQ + 255) / 256)), dim3(256), 0, s,
                       d.Mpart, p.nQ, p.K1, d.Mglob);
    a.Mglob = d.Mglob;
  }
  if (timing) (void)hipEventRecord(prof->ev[1], s);
  if (MODE == kF32Lse)
    e = launch_pass1_g<DIM, MODE, 2>(p.G, a, p.n_blocks, p.K1, s);
  else
    e = launch_pass1_g<DIM, MODE, 1>(p.G, a, p.n_blocks, p.K1, s);
  if (e != hipSuccess) return e;
  if (timing) (void)hipEventRecord(prof->ev[2], s);

  QFArgs q{};
  q.qx = d.qx_m;
  q.qw = d.qw_m;
  q.qx2 = d.qx2_m;
  q.Mpart = d.Mpart;
  q.Spart = d.Spart;
  q.Tpart = d.Tpart;
  q.cell_r2m = d.cell_r2m;
  q.qrec_r = d.qrec_r;
  q.qchunk = d.qchunk;
  q.G = p.G;
  q.Fpart = d.Fpart;
  q.Cpart = d.Cpart;
  q.Upart = d.Upart;
  q.nQ = p.nQ;
  q.nq = p.nq;
  q.K1 = p.K1;
  q.eps = p.eps;
  q.plain_coef = 0.5 / p.eps;
  q.ln_scale = ln_scale;
  hipLaunchKernelGGL((k_qfin<DIM, MODE>), dim3((unsigned)p.nFpart), dim3(256), 0, s, q);
  if (prune) {
    const int64_t ncr = grid_rows(p.cg);
    hipLaunchKernelGGL((k_cellv<DIM>), dim3(nblk256(p.n_cells)), dim3(256), 0, s, d.qrec_r,
                       p.n_cells, p.nq, theta, 1.0 / coef, MODE == kF32Lse ? 1 : 0, d.vcell);
    hipLaunchKernelGGL(k_binred<false>, dim3((unsigned)((p.n_cbins + 3) / 4)), dim3(256), 0, s,
                       d.vcell, d.cstart, p.n_cbins, d.vbin);
    hipLaunchKernelGGL(k_rowred<false>, dim3(nblk256(ncr)), dim3(256), 0, s, d.vbin, ncr,
                       p.cg.n[0], d.vrow);
  }
  if (timing) (void)hipEventRecord(prof->ev[3], s);

  GArgs g{};
  g.trec_m = d.trec_m;
  g.twbox = d.twbox;
  g.cc_r = d.cc_r;
  g.cstart = d.cstart;
  g.qrec_r = d.qrec_r;
  g.qchunk = d.qchunk;
  g.Gpart = d.Gpart;
  g.cg = p.cg;
  g.N = p.N;
  g.n_cells = p.n_cells;
  g.nq = p.nq;
  g.K2 = p.K2;
  g.dense = p.dense;
  g.smax = p.smax;
  g.Rq = p.Rq;
  g.s_in = p.s_in;
  g.Bpart = d_Bpart;
  g.prune = prune ? 1 : 0;
  g.vcell = d.vcell;
  g.vbin = d.vbin;
  g.vrow = d.vrow;
  g.rq = p.rq;
  g.inv_coef = 1.0 / coef;
  g.eps = p.eps;
  e = d_bary ? launch_grad_g<DIM, MODE, true>(p.G, g, p.n_twaves, p.K2, s)
             : launch_grad_g<DIM, MODE, false>(p.G, g, p.n_twaves, p.K2, s);
  if (e != hipSuccess) return e;
  if (timing) (void)hipEventRecord(prof->ev[4], s);
  GFArgs gf{d.Gpart, d.trec_m, d.tm2o, d_grad, p.N, p.K2};
  hipLaunchKernelGGL((k_gfin<DIM>), dim3((unsigned)((p.N + 255) / 256)), dim3(256), 0, s, gf);
  if (d_bary) {
    Origin o{{p.origin[0], p.origin[1], p.origin[2]}};
    hipLaunchKernelGGL((k_bfin<DIM>), dim3((unsigned)((p.N + 255) / 256)), dim3(256), 0, s,
                       d.Gpart, d_Bpart, d.tm2o, o, p.N, p.K2, d_bary);
  }
  FinArgs fa{d.Fpart, d.Cpart, d.Upart, d.PNpart, p.nFpart, p.nPpart, d_FC, d.ucount};
  hipLaunchKernelGGL(k_final, dim3(1), dim3(256), 0, s, fa);
  if (timing) (void)hipEventRecord(prof->ev[5], s);
  return hipGetLastError();
}

hipError_t launch_eval(const Plan &p, Dev &d, const double *d_psi, double *d_grad,
                       double *d_FC, hipStream_t s, Profile *prof) {
  if (p.dim == 2) {
    switch (p.mode) {
      case kF64Lse: return eval_t<2, kF64Lse>(p, d, d_psi, d_grad, d_FC, s, prof);
      case kF64Plain: return eval_t<2, kF64Plain>(p, d, d_psi, d_grad, d_FC, s, prof);
      case kF32Lse: return eval_t<2, kF32Lse>(p, d, d_psi, d_grad, d_FC, s, prof);
    }
  } else if (p.dim == 3) {
    switch (p.mode) {
      case kF64Lse: return eval_t<3, kF64Lse>(p, d, d_psi, d_grad, d_FC, s, prof);
      case kF64Plain: return eval_t<3, kF64Plain>(p, d, d_psi, d_grad, d_FC, s, prof);
      case kF32Lse: return eval_t<3, kF32Lse>(p, d, d_psi, d_grad, d_FC, s, prof);
    }
  }
  return hipErrorInvalidValue;
}

hipError_t launch_eval_bary(const Plan &p, Dev &d, const double *d_psi, double *d_grad,
                            double *d_FC, double *d_bary, double *d_Bpart, hipStream_t s) {
  if (p.dim == 2) {
    switch (p.mode) {
      case kF64Lse: return eval_t<2, kF64Lse>(p, d, d_psi, d_grad, d_FC, s, nullptr, d_bary, d_Bpart);
      case kF64Plain: return eval_t<2, kF64Plain>(p, d, d_psi, d_grad, d_FC, s, nullptr, d_bary, d_Bpart);
      case kF32Lse: return eval_t<2, kF32Lse>(p, d, d_psi, d_grad, d_FC, s, nullptr, d_bary, d_Bpart);
    }
  } else if (p.dim == 3) {
    switch (p.mode) {
      case kF64Lse: return eval_t<3, kF64Lse>(p, d, d_psi, d_grad, d_FC, s, nullptr, d_bary, d_Bpart);
      case kF64Plain: return eval_t<3, kF64Plain>(p, d, d_psi, d_grad, d_FC, s, nullptr, d_bary, d_Bpart);
      case kF32Lse: return eval_t<3, kF32Lse>(p, d, d_psi, d_grad, d_FC, s, nullptr, d_bary, d_Bpart);
    }
  }
  return hipErrorInvalidValue;
}

hipError_t launch_candidate_stats(const Plan &p, Dev &d, int64_t *d_count, uint64_t *d_hash,
                                  hipStream_t s) {
  CSArgs a{};
  a.cc = d.cc_m;
  a.blkbox = p.small ? d.blkbox_s : d.blkbox;
  a.trec = d.trec_r;
  a.tstart = d.tstart;
  a.tr2o = d.tr2o;
  a.cell_m2o = d.cell_m2o;
  a.count = d_count;
  a.hash = d_hash;
  a.tg = p.tg;
  a.n_groups = p.n_groups;
  a.N = p.N;
  a.gpc = p.gpc;
  a.dense = p.dense;
  a.smax = p.smax;
  a.Rq = p.Rq;
  if (p.dim == 2)
    hipLaunchKernelGGL(k_cstats<2>, dim3((unsigned)p.n_blocks), dim3(64), 0, s, a);
  else
    hipLaunchKernelGGL(k_cstats<3>, dim3((unsigned)p.n_blocks), dim3(64), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_covering(const Plan &p, Dev &d, double *d_cellmin, hipStream_t s) {
  const unsigned nb = (unsigned)((p.n_cells + 255) / 256);
  if (p.dim == 2)
    hipLaunchKernelGGL(k_cover<2>, dim3(nb), dim3(256), 0, s, d.cc_m, d.ty_m, p.n_cells, p.N,
                       d_cellmin);
  else
    hipLaunchKernelGGL(k_cover<3>, dim3(nb), dim3(256), 0, s, d.cc_m, d.ty_m, p.n_cells, p.N,
                       d_cellmin);
  return hipGetLastError();
}

}  // namespace sdot
