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
//                PASS 0: max exponent (FP64 modes),
//                PASS 1: sum_j nu_j exp(t_qj - M_q) with the global max (FP64 modes),
//                PASS 2: one online sweep (FP32 mode): per segment an integer
//                        reference value M and S = sum nu exp2(t - M); M is raised
//                        (rarely) to ceil(t) when a term would exceed it, which
//                        rescales S by an exact power of two.
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

#include "sdot_internal.hpp"

namespace sdot {

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------

template <int DIM>
__device__ __forceinline__ bool in_set(const double *y, const double *c, double smax) {
#pragma clang fp contract(off)
  double s = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const double t = y[d] - c[d];
    s = s + t * t;
  }
  return s <= smax;
}

template <int MODE>
__device__ __forceinline__ double expo(double d) {
  if constexpr (MODE == kF32Lse) {
    return (double)__builtin_amdgcn_exp2f((float)d);
  } else {
    return exp(d);
  }
}

// Visit the bin rows (x-runs of bins, contiguous in row-major bin order) whose
// bins can hold a point within Rq of the box [lo,hi]; every nseg-th visited row
// (phase seg) is handed to body(first_item, end_item).
template <int DIM, class Body>
__device__ __forceinline__ void for_rows(const Grid &g, const int32_t *__restrict__ start,
                                         const double *lo, const double *hi, double Rq,
                                         int seg, int nseg, Body &&body) {
  int b0[3] = {0, 0, 0}, b1[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    b0[d] = bin_of(lo[d] - Rq, g.o[d], g.inv_h, g.n[d]);
    b1[d] = bin_of(hi[d] + Rq, g.o[d], g.inv_h, g.n[d]);
  }
  const double R2 = Rq * Rq;
  int phase = 0;
  for (int bz = b0[2]; bz <= b1[2]; ++bz) {
    double dz = 0.0;
    if (DIM == 3) {
      const double zl = g.o[2] + bz * g.h, zh = zl + g.h;
      dz = fmax(0.0, fmax(zl - hi[2], lo[2] - zh));
    }
    for (int by = b0[1]; by <= b1[1]; ++by) {
      const double yl = g.o[1] + by * g.h, yh = yl + g.h;
      const double dy = fmax(0.0, fmax(yl - hi[1], lo[1] - yh));
      const double r2 = R2 - dy * dy - dz * dz;
      if (r2 < 0.0) continue;
      const int mine = (phase == seg);
      phase = (phase + 1 == nseg) ? 0 : phase + 1;
      if (!mine) continue;
      const double wx = sqrt(r2);
      const int x0 = bin_of(lo[0] - wx, g.o[0], g.inv_h, g.n[0]);
      const int x1 = bin_of(hi[0] + wx, g.o[0], g.inv_h, g.n[0]);
      const int base = (bz * g.n[1] + by) * g.n[0];
      const int s = __builtin_amdgcn_readfirstlane(start[base + x0]);
      const int e = __builtin_amdgcn_readfirstlane(start[base + x1 + 1]);
      if (s < e) body(s, e);
    }
  }
}

// Fixed-pattern block sum / max (blockDim.x == 256). Result valid in all threads.
__device__ __forceinline__ double block_sum_256(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}
__device__ __forceinline__ double block_max_256(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Target record (64 bytes):
//   [0]=a  [1..DIM]=y'/eps  [1+DIM..2DIM]=y raw  [1+2DIM]=nu
template <int DIM>
struct TRec {
  double a, yl[DIM], y[DIM], nu;
  __device__ __forceinline__ void load(const double *__restrict__ r) {
    a = r[0];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      yl[d] = r[1 + d];
      y[d] = r[1 + DIM + d];
    }
    nu = r[1 + 2 * DIM];
  }
};

// ---------------------------------------------------------------------------
// k_prep: target records
// ---------------------------------------------------------------------------
struct PrepArgs {
  const double *y;     // [N*dim] caller order
  const double *nu;    // [N]
  const double *psi;   // [N]
  const int32_t *pos_r, *pos_m;
  double *trec_r, *trec_m;
  double *PNpart;
  double *psi_prev;    // [N] psi of the previous FP32 evaluation (history mode)
  double *DMpart, *Dmpart;  // block max / min of psi - psi_prev
  int64_t N;
  int hist;
  double o[3];
  double coef;         // inv_eps * log-base scale
};

template <int DIM>
__global__ __launch_bounds__(256) void k_prep(PrepArgs a) {
  __shared__ double sh[256];
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double pn = 0.0, dmx = -INFINITY, dmn = INFINITY;
  if (j < a.N) {
    double yr[DIM], yp[DIM];
    double y2 = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      yr[d] = a.y[j * DIM + d];
      yp[d] = yr[d] - a.o[d];
      y2 += yp[d] * yp[d];
    }
    const double psi = a.psi[j], nu = a.nu[j];
    double rec[kRec];
    rec[0] = (psi - 0.5 * y2) * a.coef;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      rec[1 + d] = yp[d] * a.coef;
      rec[1 + DIM + d] = yr[d];
    }
    rec[1 + 2 * DIM] = nu;
#pragma unroll
    for (int k = 2 + 2 * DIM; k < kRec; ++k) rec[k] = 0.0;
    double *r1 = a.trec_r + (int64_t)a.pos_r[j] * kRec;
    double *r2 = a.trec_m + (int64_t)a.pos_m[j] * kRec;
#pragma unroll
    for (int k = 0; k < kRec; ++k) {
      r1[k] = rec[k];
      r2[k] = rec[k];
    }
    pn = psi * nu;
    if (a.psi_prev) {
      if (a.hist) {
        const double dd = psi - a.psi_prev[j];
        dmx = dd;
        dmn = dd;
      }
      a.psi_prev[j] = psi;
    }
  }
  const double s = block_sum_256(pn, sh);
  if (threadIdx.x == 0) a.PNpart[blockIdx.x] = s;
  if (a.psi_prev) {
    const double m1 = block_max_256(dmx, sh);
    const double m2 = -block_max_256(-dmn, sh);
    if (threadIdx.x == 0) {
      a.DMpart[blockIdx.x] = m1;
      a.Dmpart[blockIdx.x] = m2;
    }
  }
}

// Global max / min of the psi change (history mode): dstat[0] = max, [1] = min.
__global__ __launch_bounds__(256) void k_dred(const double *DMpart, const double *Dmpart,
                                              int64_t n, double *dstat) {
  __shared__ double sh[256];
  double mx = -INFINITY, mn = INFINITY;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    mx = fmax(mx, DMpart[i]);
    mn = fmin(mn, Dmpart[i]);
  }
  const double a = block_max_256(mx, sh);
  const double b = -block_max_256(-mn, sh);
  if (threadIdx.x == 0) {
    dstat[0] = a;
    dstat[1] = b;
  }
}

// ---------------------------------------------------------------------------
// k_pass1: q-major sweeps
// ---------------------------------------------------------------------------
struct P1Args {
  const double *qx;      // [nQ*DIM] Morton order, shifted
  const double *qx2;     // [nQ]
  const double *cc;      // [n_cells*DIM] raw centers, Morton order
  const double *blkbox;  // [n_blocks*2*DIM]
  const double *trec;    // [N*kRec] row order
  const int32_t *tstart;
  const double *Mpart;   // [K1*nQ]  read by PASS 1
  const double *Mhist;   // [nQ]     PASS 2: M of the previous evaluation
  const double *dstat;   // PASS 2: max / min psi change since then
  double *out;           // [K1*nQ]  M (PASS 0/2) or S (PASS 1)
  double *out2;          // [K1*nQ]  S (PASS 2)
  Grid tg;
  int64_t n_groups, nQ, N;
  int nq, gpc, K1, dense, hist;
  double smax, Rq, plain_coef, coef;
};

template <int DIM, int G, int MODE, int PASS>
__global__ __launch_bounds__(64) void k_pass1(P1Args a) {
  const int lane = threadIdx.x;
  const int seg = blockIdx.y;
  const int64_t grp = (int64_t)blockIdx.x * kWave + lane;
  const bool active = grp < a.n_groups;
  const int64_t gg = active ? grp : 0;
  const int64_t cell = gg / a.gpc;
  const int64_t q0 = cell * a.nq + (gg - cell * a.gpc) * G;

  double c[DIM], x[G][DIM], M[G], S[G];
#pragma unroll
  for (int d = 0; d < DIM; ++d) c[d] = a.cc[cell * DIM + d];
  // history seed of the online sweep: the previous evaluation's M_q raised by
  // the largest psi increase bounds every new exponent of the targets that were
  // candidates before; skipped when the psi change spread could underflow
  bool seed = false;
  double shift = 0.0;
  if (PASS == 2 && a.hist) {
    const double dmx = a.dstat[0], dmn = a.dstat[1];
    seed = (dmx - dmn) * a.coef < 100.0;
    shift = dmx * a.coef;
  }
#pragma unroll
  for (int k = 0; k < G; ++k) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) x[k][d] = a.qx[(q0 + k) * DIM + d];
    S[k] = 0.0;
    if (PASS == 0) {
      M[k] = -INFINITY;
    } else if (PASS == 2) {
      M[k] = seed ? ceil(a.Mhist[q0 + k] + shift) : -INFINITY;
      if (!(M[k] < INFINITY)) M[k] = -INFINITY;
    } else if (MODE == kF64Plain) {
      M[k] = a.plain_coef * a.qx2[q0 + k];
    } else {
      double m = -INFINITY;
      for (int s = 0; s < a.K1; ++s) m = fmax(m, a.Mpart[(int64_t)s * a.nQ + q0 + k]);
      M[k] = m;
    }
  }
  double lo[DIM], hi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    lo[d] = a.blkbox[(int64_t)blockIdx.x * 2 * DIM + d];
    hi[d] = a.blkbox[(int64_t)blockIdx.x * 2 * DIM + DIM + d];
  }
  auto one = [&](const TRec<DIM> &r) {
    if (!a.dense && !in_set<DIM>(r.y, c, a.smax)) return;
    if (PASS == 0) {
#pragma unroll
      for (int k = 0; k < G; ++k) {
        double t = r.a;
#pragma unroll
        for (int d = 0; d < DIM; ++d) t = fma(x[k][d], r.yl[d], t);
        M[k] = fmax(M[k], t);
      }
    } else if (PASS == 1) {
#pragma unroll
      for (int k = 0; k < G; ++k) {
        double t = r.a;
#pragma unroll
        for (int d = 0; d < DIM; ++d) t = fma(x[k][d], r.yl[d], t);
        S[k] = fma(r.nu, expo<MODE>(t - M[k]), S[k]);
      }
    } else {
#pragma unroll
      for (int k = 0; k < G; ++k) {
        double t = r.a;
#pragma unroll
        for (int d = 0; d < DIM; ++d) t = fma(x[k][d], r.yl[d], t);
        double dlt = t - M[k];
        if (dlt > 0.0) {
          // rare: raise the reference value to the next integer, so the
          // rescale of the running sum is an exact power of two
          const double Mn = ceil(t);
          S[k] = (M[k] > -INFINITY) ? ldexp(S[k], (int)(M[k] - Mn)) : 0.0;
          M[k] = Mn;
          dlt = t - Mn;
        }
        S[k] = fma(r.nu, (double)__builtin_amdgcn_exp2f((float)dlt), S[k]);
      }
    }
  };
  auto body = [&](int j0, int j1) {
    // records are fetched two at a time ahead of their use
    for (int j = j0; j < j1; j += 2) {
      const int j2 = (j + 1 < j1) ? j + 1 : j;
      TRec<DIM> r0, r1;
      r0.load(a.trec + (int64_t)j * kRec);
      r1.load(a.trec + (int64_t)j2 * kRec);
      one(r0);
      if (j + 1 < j1) one(r1);
    }
  };

  if (a.dense) {
    const int64_t j0 = a.N * seg / a.K1, j1 = a.N * (seg + 1) / a.K1;
    body((int)j0, (int)j1);
  } else {
    for_rows<DIM>(a.tg, a.tstart, lo, hi, a.Rq, seg, a.K1, body);
  }
  if (active) {
#pragma unroll
    for (int k = 0; k < G; ++k) {
      const int64_t o = (int64_t)seg * a.nQ + q0 + k;
      if (PASS == 0) a.out[o] = M[k];
      if (PASS == 1) a.out[o] = S[k];
      if (PASS == 2) {
        a.out[o] = M[k];
        a.out2[o] = S[k];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// k_qfin: per-q merge, functional, scale, q records (row cell order)
// ---------------------------------------------------------------------------
struct QFArgs {
  const double *qx, *qw, *qx2, *Mpart, *Spart;
  const int64_t *cell_r2m;
  double *qrec_r, *Fpart, *Cpart, *Upart, *Mhist;
  int64_t nQ;
  int nq, K1;
  double eps, plain_coef, ln_scale;
};

template <int DIM, int MODE>
__global__ __launch_bounds__(256) void k_qfin(QFArgs a) {
  __shared__ double sh[256];
  const int64_t qr = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double Fq = 0.0, Cq = 0.0, Uq = 0.0;
  if (qr < a.nQ) {
    const int64_t cr = qr / a.nq, k = qr - cr * a.nq;
    const int64_t qm = a.cell_r2m[cr] * a.nq + k;
    double x[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) x[d] = a.qx[qm * DIM + d];
    const double x2 = a.qx2[qm];
    double M, S = 0.0;
    if (MODE == kF64Plain) {
      M = a.plain_coef * x2;
      for (int s = 0; s < a.K1; ++s) S += a.Spart[(int64_t)s * a.nQ + qm];
    } else if (MODE == kF64Lse) {
      M = -INFINITY;
      for (int s = 0; s < a.K1; ++s) M = fmax(M, a.Mpart[(int64_t)s * a.nQ + qm]);
      for (int s = 0; s < a.K1; ++s) S += a.Spart[(int64_t)s * a.nQ + qm];
    } else {
      // segments carry their own reference values: S = sum S_s 2^(M_s - M)
      M = -INFINITY;
      for (int s = 0; s < a.K1; ++s) M = fmax(M, a.Mpart[(int64_t)s * a.nQ + qm]);
      if (M > -INFINITY)
        for (int s = 0; s < a.K1; ++s) {
          const double ss = a.Spart[(int64_t)s * a.nQ + qm];
          if (ss > 0.0) S += ldexp(ss, (int)(a.Mpart[(int64_t)s * a.nQ + qm] - M));
        }
      a.Mhist[qm] = M;
      // every term far below the reference value: the FP32 exponentials lost
      // their range; the host repeats the evaluation without the seed
      if (M > -INFINITY && S < 1e-40) Uq = 1.0;
    }
    const double w = a.qw[qm];
    double Mrec = INFINITY, scale = 0.0;
    // q-points whose sum vanished are skipped (SotSolver.cc:854)
    if (S > 0.0 && M > -INFINITY) {
      Fq = w * (a.eps * (M * a.ln_scale + log(S)) - 0.5 * x2);
      scale = w / S;
      Cq = scale;
      Mrec = M;
    }
    double *r = a.qrec_r + qr * kRec;
#pragma unroll
    for (int d = 0; d < DIM; ++d) r[d] = x[d];
    r[DIM] = Mrec;
    r[DIM + 1] = scale;
    r[DIM + 2] = (scale > 0.0) ? S : 0.0;
#pragma unroll
    for (int t = DIM + 3; t < kRec; ++t) r[t] = 0.0;
  }
  const double fs = block_sum_256(Fq, sh);
  const double cs = block_sum_256(Cq, sh);
  const double us = block_sum_256(Uq, sh);
  if (threadIdx.x == 0) {
    a.Fpart[blockIdx.x] = fs;
    a.Cpart[blockIdx.x] = cs;
    a.Upart[blockIdx.x] = us;
  }
}

// ---------------------------------------------------------------------------
// k_grad: target-major gradient partials
// ---------------------------------------------------------------------------
struct GArgs {
  const double *trec_m;  // [N*kRec] Morton order
  const double *twbox;   // [n_twaves*2*DIM]
  const double *cc_r;    // [n_cells*4] raw centers, row order
  const int32_t *cstart;
  const double *qrec_r;  // [nQ*kRec]
  double *Gpart;         // [K2*N]
  Grid cg;
  int64_t N, n_cells;
  int nq, K2, dense;
  double smax, Rq;
};

template <int DIM, int MODE, int G>
__global__ __launch_bounds__(64) void k_grad(GArgs a) {
  const int seg = blockIdx.y;
  const int64_t j = (int64_t)blockIdx.x * kWave + threadIdx.x;
  const bool active = j < a.N;
  TRec<DIM> me;
  me.load(a.trec_m + (active ? j : 0) * kRec);
  double lo[DIM], hi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    lo[d] = a.twbox[(int64_t)blockIdx.x * 2 * DIM + d];
    hi[d] = a.twbox[(int64_t)blockIdx.x * 2 * DIM + DIM + d];
  }
  double acc = 0.0;
  const int nq = a.nq;
  auto chunk = [&](const double *__restrict__ qp) {
    double qx[G][DIM], qm[G], qs[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) qx[k][d] = qp[k * kRec + d];
      qm[k] = qp[k * kRec + DIM];
      qs[k] = qp[k * kRec + DIM + 1];
    }
#pragma unroll
    for (int k = 0; k < G; ++k) {
      double t = me.a;
#pragma unroll
      for (int d = 0; d < DIM; ++d) t = fma(qx[k][d], me.yl[d], t);
      const double e = expo<MODE>(t - qm[k]);
      if (MODE == kF64Plain) {
        // the reference adds only exp_terms > 0 (SotSolver.cc:292); in plain
        // mode w/S_q can overflow to inf when S_q is subnormal
        if (e > 0.0) acc = fma(qs[k], e, acc);
      } else {
        acc = fma(qs[k], e, acc);
      }
    }
  };
  auto body = [&](int c0, int c1) {
    for (int c = c0; c < c1; ++c) {
      if (!a.dense) {
        const double *__restrict__ cp = a.cc_r + (int64_t)c * 4;
        double cv[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) cv[d] = cp[d];
        if (!in_set<DIM>(me.y, cv, a.smax)) continue;
      }
      const double *qp = a.qrec_r + (int64_t)c * nq * kRec;
      for (int k0 = 0; k0 < nq; k0 += G) chunk(qp + (int64_t)k0 * kRec);
    }
  };
  if (a.dense) {
    const int64_t c0 = a.n_cells * seg / a.K2, c1 = a.n_cells * (seg + 1) / a.K2;
    body((int)c0, (int)c1);
  } else {
    for_rows<DIM>(a.cg, a.cstart, lo, hi, a.Rq, seg, a.K2, body);
  }
  if (active) a.Gpart[(int64_t)seg * a.N + j] = acc;
}

// ---------------------------------------------------------------------------
// k_gfin: g_j = nu_j * sum_seg G - nu_j, scattered to caller order
// ---------------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(256) void k_gfin(const double *Gpart, const double *trec_m,
                                              const int32_t *tm2o, double *grad, int64_t N,
                                              int K2) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= N) return;
  double s = 0.0;
  for (int k = 0; k < K2; ++k) s += Gpart[(int64_t)k * N + j];
  const double nu = trec_m[j * kRec + 1 + 2 * DIM];
  grad[tm2o[j]] = nu * s - nu;
}

// Fixed-order totals: out[0] = sum F_q - sum psi.nu, out[1] = sum C_q,
// *ucount = number of q-points whose FP32 sum lost its range.
__global__ __launch_bounds__(256) void k_final(const double *Fpart, const double *Cpart,
                                               const double *Upart, int64_t nF,
                                               const double *PNpart, int64_t nP, double *out,
                                               double *ucount) {
  __shared__ double sh[256];
  double f = 0.0, c = 0.0, u = 0.0, p = 0.0;
  for (int64_t i = threadIdx.x; i < nF; i += 256) {
    f += Fpart[i];
    c += Cpart[i];
    u += Upart[i];
  }
  for (int64_t i = threadIdx.x; i < nP; i += 256) p += PNpart[i];
  const double fs = block_sum_256(f, sh);
  const double cs = block_sum_256(c, sh);
  const double us = block_sum_256(u, sh);
  const double ps = block_sum_256(p, sh);
  if (threadIdx.x == 0) {
    out[0] = fs - ps;
    out[1] = cs;
    *ucount = us;
  }
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
  auto body = [&](int j0, int j1) {
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
    body(0, (int)a.N);
  else
    for_rows<DIM>(a.tg, a.tstart, lo, hi, a.Rq, 0, 1, body);
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

// A[j][k] += sum_q P[q][j] P[q][k] on 64x64 blocks with bk >= bj.  Four waves
// per block, each a 32x32 quadrant as 2x2 MFMA 16x16 tiles.
__global__ __launch_bounds__(256) void k_syrk(const double *__restrict__ P, int64_t rows,
                                              int64_t Npad, double *__restrict__ A) {
  const int bj = blockIdx.y, bk = blockIdx.x;
  if (bk < bj) return;
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int wj = w >> 1, wk = w & 1;
  const int64_t j0 = (int64_t)bj * 64 + wj * 32, k0 = (int64_t)bk * 64 + wk * 32;
  sdot_d4 acc[2][2];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int u = 0; u < 2; ++u) acc[t][u] = sdot_d4{0.0, 0.0, 0.0, 0.0};
  const int lr = l >> 4, lc = l & 15;
  for (int64_t q0 = 0; q0 < rows; q0 += 4) {
    const double *row = P + (q0 + lr) * Npad;
    double a[2], b[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      a[t] = row[j0 + t * 16 + lc];
      b[t] = row[k0 + t * 16 + lc];
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int u = 0; u < 2; ++u)
        acc[t][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b[u], acc[t][u], 0, 0, 0);
  }
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        // f64 16x16x4 output: register i of lane l holds row 4*i + l/16, column l%16
        const int64_t jj = j0 + t * 16 + i * 4 + lr, kk = k0 + u * 16 + lc;
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
  const unsigned nb = (unsigned)(Npad / 64);
  hipLaunchKernelGGL(k_syrk, dim3(nb, nb), dim3(256), 0, s, d_P, rows, Npad, d_A);
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
  hipLaunchKernelGGL((k_pass1<DIM, G, MODE, PASS>), dim3((unsigned)nb, (unsigned)K1), dim3(64),
                     0, s, a);
}

template <int DIM, int MODE, int PASS>
static hipError_t launch_pass1_g(int G, const P1Args &a, int64_t nb, int K1, hipStream_t s) {
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

template <int DIM, int MODE>
static void launch_grad_g(int G, const GArgs &g, int64_t ntw, int K2, hipStream_t s) {
  const dim3 gg((unsigned)ntw, (unsigned)K2);
  switch (G) {
    case 1: hipLaunchKernelGGL((k_grad<DIM, MODE, 1>), gg, dim3(64), 0, s, g); break;
    case 2: hipLaunchKernelGGL((k_grad<DIM, MODE, 2>), gg, dim3(64), 0, s, g); break;
    case 3: hipLaunchKernelGGL((k_grad<DIM, MODE, 3>), gg, dim3(64), 0, s, g); break;
    case 4: hipLaunchKernelGGL((k_grad<DIM, MODE, 4>), gg, dim3(64), 0, s, g); break;
    case 5: hipLaunchKernelGGL((k_grad<DIM, MODE, 5>), gg, dim3(64), 0, s, g); break;
    case 6: hipLaunchKernelGGL((k_grad<DIM, MODE, 6>), gg, dim3(64), 0, s, g); break;
    case 7: hipLaunchKernelGGL((k_grad<DIM, MODE, 7>), gg, dim3(64), 0, s, g); break;
    case 8: hipLaunchKernelGGL((k_grad<DIM, MODE, 8>), gg, dim3(64), 0, s, g); break;
    default: hipLaunchKernelGGL((k_grad<DIM, MODE, 9>), gg, dim3(64), 0, s, g); break;
  }
}

template <int DIM, int MODE>
static hipError_t eval_t(const Plan &p, Dev &d, const double *d_psi, double *d_grad,
                         double *d_FC, hipStream_t s, Profile *prof) {
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
  pa.DMpart = d.DMpart;
  pa.Dmpart = d.Dmpart;
  pa.N = p.N;
  pa.hist = p.hist;
  for (int k = 0; k < 3; ++k) pa.o[k] = p.origin[k];
  pa.coef = coef;
  hipLaunchKernelGGL((k_prep<DIM>), dim3((unsigned)p.nPpart), dim3(256), 0, s, pa);
  if (MODE == kF32Lse && p.hist)
    hipLaunchKernelGGL(k_dred, dim3(1), dim3(256), 0, s, d.DMpart, d.Dmpart, p.nPpart, d.dstat);

  P1Args a{};
  a.qx = d.qx_m;
  a.qx2 = d.qx2_m;
  a.cc = d.cc_m;
  a.blkbox = d.blkbox;
  a.trec = d.trec_r;
  a.tstart = d.tstart;
  a.Mpart = d.Mpart;
  a.Mhist = d.Mhist_m;
  a.dstat = d.dstat;
  a.tg = p.tg;
  a.n_groups = p.n_groups;
  a.nQ = p.nQ;
  a.N = p.N;
  a.nq = p.nq;
  a.gpc = p.gpc;
  a.K1 = p.K1;
  a.dense = p.dense;
  a.hist = p.hist;
  a.smax = p.smax;
  a.Rq = p.Rq;
  a.plain_coef = 0.5 / p.eps;
  a.coef = coef;
  hipError_t e;
  if (MODE == kF64Lse) {
    a.out = d.Mpart;
    e = launch_pass1_g<DIM, MODE, 0>(p.G, a, p.n_blocks, p.K1, s);
    if (e != hipSuccess) return e;
  }
  if (timing) (void)hipEventRecord(prof->ev[1], s);
  if (MODE == kF32Lse) {
    a.out = d.Mpart;
    a.out2 = d.Spart;
    e = launch_pass1_g<DIM, MODE, 2>(p.G, a, p.n_blocks, p.K1, s);
  } else {
    a.out = d.Spart;
    e = launch_pass1_g<DIM, MODE, 1>(p.G, a, p.n_blocks, p.K1, s);
  }
  if (e != hipSuccess) return e;
  if (timing) (void)hipEventRecord(prof->ev[2], s);

  QFArgs q{};
  q.qx = d.qx_m;
  q.qw = d.qw_m;
  q.qx2 = d.qx2_m;
  q.Mpart = d.Mpart;
  q.Spart = d.Spart;
  q.cell_r2m = d.cell_r2m;
  q.qrec_r = d.qrec_r;
  q.Fpart = d.Fpart;
  q.Cpart = d.Cpart;
  q.Upart = d.Upart;
  q.Mhist = d.Mhist_m;
  q.nQ = p.nQ;
  q.nq = p.nq;
  q.K1 = p.K1;
  q.eps = p.eps;
  q.plain_coef = 0.5 / p.eps;
  q.ln_scale = ln_scale;
  hipLaunchKernelGGL((k_qfin<DIM, MODE>), dim3((unsigned)p.nFpart), dim3(256), 0, s, q);
  if (timing) (void)hipEventRecord(prof->ev[3], s);

  GArgs g{};
  g.trec_m = d.trec_m;
  g.twbox = d.twbox;
  g.cc_r = d.cc_r;
  g.cstart = d.cstart;
  g.qrec_r = d.qrec_r;
  g.Gpart = d.Gpart;
  g.cg = p.cg;
  g.N = p.N;
  g.n_cells = p.n_cells;
  g.nq = p.nq;
  g.K2 = p.K2;
  g.dense = p.dense;
  g.smax = p.smax;
  g.Rq = p.Rq;
  launch_grad_g<DIM, MODE>(p.G, g, p.n_twaves, p.K2, s);
  if (timing) (void)hipEventRecord(prof->ev[4], s);
  hipLaunchKernelGGL((k_gfin<DIM>), dim3((unsigned)((p.N + 255) / 256)), dim3(256), 0, s,
                     d.Gpart, d.trec_m, d.tm2o, d_grad, p.N, p.K2);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(256), 0, s, d.Fpart, d.Cpart, d.Upart, p.nFpart,
                     d.PNpart, p.nPpart, d_FC, d.dstat + 2);
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

hipError_t launch_candidate_stats(const Plan &p, Dev &d, int64_t *d_count, uint64_t *d_hash,
                                  hipStream_t s) {
  CSArgs a{};
  a.cc = d.cc_m;
  a.blkbox = d.blkbox;
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
