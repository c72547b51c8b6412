// sdot_points.hip -- gfx950 kernels of the point queries made after a solve,
// FP64 throughout:
//
//   conditional densities (SotSolver::get_potential_conditioned_density,
//   SotSolver.cc:973-1046): at a point p, the targets j with |p - y_j| <= R,
//     e_pj = (psi_j - pow(|p - y_j|, 2) / 2) * (1/eps),  M_p = max_j e_pj,
//     S_p  = sum_j nu_j exp(e_pj - M_p),
//     c_i(p) = rho_p exp(e_pi - M_p) / S_p   (0 where i is not a candidate
//                                             of p or S_p <= 0);
//   transport maps (ModalStrategy / BarycentricStrategy::compute_map,
//   OptimalTransportPlan.cc:99-189, 218-310): the targets with
//   |x - y_j| < radius (every target when radius <= 0; the nearest target
//   when there is none),
//     modal:       y of the maximum of psi_j - pow(|x - y_j|, 2) / 2 + eps log nu_j
//                  (the first maximum in index order; -inf where nu_j <= 0),
//     barycentric: sum_j nu_j e^(e_j - M) y_j / sum_j nu_j e^(e_j - M) with
//                  e_j = (psi_j - pow(|x - y_j|, 2) / 2) / eps (the nearest
//                  target where the sum is not positive).
//
//   k_pt_sweep<DIM, OP>  lane = query point (waves of 64 Morton-ordered points),
//                        grid.y = K segments of the target bin rows (the
//                        row split of for_rows); per segment partials, no atomics:
//                          OP 0: maximum exponent and candidate count,
//                          OP 1: sum (and first moments for the barycentric map)
//                                against the maximum over every segment,
//                          OP 2: modal maximum (score, caller index) and count.
//   k_pt_fin<DIM, OP>    per point: segments merged in fixed order, results
//                        written in the caller's order; points that need the
//                        nearest-target fallback are flagged.
//   k_pt_nearest<DIM>    the flagged points: nearest target (lowest index
//                        among equal distances) by scans of growing radius.
//   k_pt_cond            per (requested index, point): c_i(p).
//
// Distances are the reference's: s = ((0 + t0^2) + t1^2) + t2^2 without FMA
// contraction, d = sqrt(s) (correctly rounded), pow(d, 2) = d * d.  Candidate
// tests run on s against host-computed bounds: s <= smax (largest s with
// sqrt(s) <= R, or with sqrt(s) < radius for the strict test); sqrt is
// monotone, so the tests are exact.  The modal log terms eps log nu_j come
// from the host (the same libm call as the reference's).
#include <hip/hip_runtime.h>

#include "sdot_device.hpp"

namespace sdot {

namespace {

// Exponent of the pair: (psi - sq/2) * eps_inv (conditional densities) or
// (psi - sq/2) / eps (barycentric map), as the two reference functions write it.
__device__ __forceinline__ double pt_exponent(const PtArgs &a, double psi, double sq) {
  const double u = psi - 0.5 * sq;
  return a.divide ? u / a.eps : u * a.eps_inv;
}

// (score, index) order of the modal map: the larger score; among equal finite
// scores the lower caller index; a point with no finite score keeps index 0.
__device__ __forceinline__ void modal_merge(double &best, int64_t &bj, double s, int64_t j) {
  if (s > best || (s == best && s > -INFINITY && j < bj)) {
    best = s;
    bj = j;
  }
}

}  // namespace

template <int DIM, int OP>
__global__ __launch_bounds__(64) void k_pt_sweep(PtArgs a) {
  const int64_t p = (int64_t)blockIdx.x * kWave + threadIdx.x;
  const bool active = p < a.n;
  const int64_t pp = active ? p : a.n - 1;
  const int seg = blockIdx.y;
  double x[DIM], lo[DIM], hi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) x[d] = a.px[pp * DIM + d];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    lo[d] = wave_min(x[d]);
    hi[d] = wave_max(x[d]);
  }
  double M = -INFINITY, S = 0.0, B[DIM], best = -INFINITY;
  int64_t bj = 0, cnt = 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) B[d] = 0.0;
  if (OP == 1)
    for (int s = 0; s < a.K; ++s) M = fmax(M, a.Mpart[(int64_t)s * a.n + pp]);
  cdouble *rec = cptr(a.rec);
  auto body = [&](int j0, int j1, auto) {
    for (int j = j0; j < j1; ++j) {
      cdouble *r = rec + (int64_t)j * kRec;
      double y[DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) y[d] = r[d];
      const double s = sq_dist<DIM>(x, y);
      if (!(s <= a.smax)) continue;
      const double dd = sqrt(s);
      const double sq = dd * dd;
      const double psi = r[3];
      if (OP == 0) {
        M = fmax(M, pt_exponent(a, psi, sq));
        ++cnt;
      } else if (OP == 1) {
        const double w = r[4] * exp(pt_exponent(a, psi, sq) - M);
        S += w;
#pragma unroll
        for (int d = 0; d < DIM; ++d) B[d] += w * y[d];
      } else {
        modal_merge(best, bj, psi - 0.5 * sq + r[5], (int64_t)r[6]);
        ++cnt;
      }
    }
  };
  for_rows<DIM>(a.tg, a.tstart, lo, hi, a.Rq, 0.0, seg, a.K, body);
  if (!active) return;
  const int64_t o = (int64_t)seg * a.n + p;
  if (OP == 0) {
    a.Mpart[o] = M;
    a.Cpart[o] = cnt;
  } else if (OP == 1) {
    a.Spart[o] = S;
    if (a.moments)
#pragma unroll
      for (int d = 0; d < DIM; ++d) a.Bpart[((int64_t)seg * DIM + d) * a.n + p] = B[d];
  } else {
    a.Mpart[o] = best;
    a.Cpart[o] = cnt;
    a.Ipart[o] = bj;
  }
}

// OP 0: conditional densities (M_p, S_p in caller order); 1: barycentric map;
// 2: modal map.  need[m] = 1 where the nearest-target fallback applies.
template <int DIM, int OP>
__global__ __launch_bounds__(256) void k_pt_fin(PtArgs a) {
  const int64_t m = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (m >= a.n) return;
  const int64_t o = a.pm2o[m];
  double M = -INFINITY, S = 0.0, B[DIM], best = -INFINITY;
  int64_t cnt = 0, bj = 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) B[d] = 0.0;
  for (int s = 0; s < a.K; ++s) {
    const int64_t i = (int64_t)s * a.n + m;
    cnt += a.Cpart[i];
    if (OP == 2) {
      modal_merge(best, bj, a.Mpart[i], a.Ipart[i]);
    } else {
      M = fmax(M, a.Mpart[i]);
      S += a.Spart[i];
      if (OP == 1)
#pragma unroll
        for (int d = 0; d < DIM; ++d) B[d] += a.Bpart[((int64_t)s * DIM + d) * a.n + m];
    }
  }
  if (OP == 0) {
    a.Mo[o] = M;
    a.So[o] = S;
    return;
  }
  const bool fallback = (a.truncated && cnt == 0) || (OP == 1 && !(S > 0.0));
  a.need[m] = fallback ? 1 : 0;
  if (fallback) return;
  double *out = a.mapped + o * DIM;
  if (OP == 1) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) out[d] = B[d] / S;
  } else {
#pragma unroll
    for (int d = 0; d < DIM; ++d) out[d] = a.ty[bj * DIM + d];
  }
}

// Nearest target of the flagged points: scans of the bins within rho of the
// flagged lanes' box, rho doubling from rho0; a lane is done once its best
// distance is at most rho (every target within rho has been seen), and every
// lane is done once rho reaches rho_all (every bin is within reach).
template <int DIM>
__global__ __launch_bounds__(64) void k_pt_nearest(PtArgs a) {
  const int64_t p = (int64_t)blockIdx.x * kWave + threadIdx.x;
  const bool mine = p < a.n && a.need[p] != 0;
  if (!__any(mine)) return;
  const int64_t pp = p < a.n ? p : a.n - 1;
  double x[DIM], lo[DIM], hi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) x[d] = a.px[pp * DIM + d];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    lo[d] = wave_min(mine ? x[d] : INFINITY);
    hi[d] = wave_max(mine ? x[d] : -INFINITY);
  }
  double bd = INFINITY;
  int64_t bj = -1;
  cdouble *rec = cptr(a.rec);
  auto body = [&](int j0, int j1, auto) {
    for (int j = j0; j < j1; ++j) {
      cdouble *r = rec + (int64_t)j * kRec;
      double y[DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) y[d] = r[d];
      const double dd = sqrt(sq_dist<DIM>(x, y));
      const int64_t jj = (int64_t)r[6];
      if (bj < 0 || dd < bd || (dd == bd && jj < bj)) {
        bd = dd;
        bj = jj;
      }
    }
  };
  for (double rho = a.rho0;; rho *= 2.0) {
    for_rows<DIM>(a.tg, a.tstart, lo, hi, rho, 0.0, 0, 1, body);
    const bool done = !mine || rho >= a.rho_all || (bj >= 0 && bd <= rho);
    if (__all(done)) break;
  }
  if (mine) {
    double *out = a.mapped + a.pm2o[p] * DIM;
#pragma unroll
    for (int d = 0; d < DIM; ++d) out[d] = a.ty[bj * DIM + d];
  }
}

// c_i(p) for every requested index (row k) and point (column p), caller order.
template <int DIM>
__global__ __launch_bounds__(256) void k_pt_cond(PtArgs a) {
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= a.n * a.n_idx) return;
  const int64_t k = t / a.n, p = t - k * a.n;
  const int64_t i = a.idx[k];
  double v = 0.0;
  const double S = a.So[p];
  if (i >= 0 && i < a.N && S > 0.0) {
    double x[DIM], y[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      x[d] = a.praw[p * DIM + d];
      y[d] = a.ty[i * DIM + d];
    }
    const double s = sq_dist<DIM>(x, y);
    if (s <= a.smax) {
      const double dd = sqrt(s);
      v = a.rho[p] * (exp(pt_exponent(a, a.psi[i], dd * dd) - a.Mo[p]) / S);
    }
  }
  a.cond[t] = v;
}

// Target records in row (bin) order: y[DIM] at 0, psi at 3, nu at 4,
// eps log nu (modal map) at 5, the caller index at 6.
template <int DIM>
__global__ __launch_bounds__(256) void k_pt_rec(PtArgs a) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= a.N) return;
  double *r = a.rec_w + (int64_t)a.tpos_r[j] * kRec;
#pragma unroll
  for (int d = 0; d < 3; ++d) r[d] = d < DIM ? a.ty[j * DIM + d] : 0.0;
  r[3] = a.psi[j];
  r[4] = a.nu[j];
  r[5] = a.logterm ? a.logterm[j] : 0.0;
  r[6] = (double)j;
  r[7] = 0.0;
}

template <int DIM>
static hipError_t points_t(PtArgs a, int op, hipStream_t s) {
  const unsigned nb = (unsigned)((a.N + 255) / 256);
  hipLaunchKernelGGL((k_pt_rec<DIM>), dim3(nb), dim3(256), 0, s, a);
  a.rec = a.rec_w;
  const int64_t nw = (a.n + kWave - 1) / kWave;
  const dim3 gs((unsigned)nw, (unsigned)a.K);
  const unsigned nf = (unsigned)((a.n + 255) / 256);
  if (op == 2) {
    hipLaunchKernelGGL((k_pt_sweep<DIM, 2>), gs, dim3(64), 0, s, a);
    hipLaunchKernelGGL((k_pt_fin<DIM, 2>), dim3(nf), dim3(256), 0, s, a);
  } else {
    hipLaunchKernelGGL((k_pt_sweep<DIM, 0>), gs, dim3(64), 0, s, a);
    hipLaunchKernelGGL((k_pt_sweep<DIM, 1>), gs, dim3(64), 0, s, a);
    if (op == 0)
      hipLaunchKernelGGL((k_pt_fin<DIM, 0>), dim3(nf), dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((k_pt_fin<DIM, 1>), dim3(nf), dim3(256), 0, s, a);
  }
  if (op == 0) {
    if (a.n_idx > 0)
      hipLaunchKernelGGL((k_pt_cond<DIM>), dim3((unsigned)((a.n * a.n_idx + 255) / 256)),
                         dim3(256), 0, s, a);
  } else {
    hipLaunchKernelGGL((k_pt_nearest<DIM>), dim3((unsigned)nw), dim3(64), 0, s, a);
  }
  return hipGetLastError();
}

hipError_t launch_points(int dim, const PtArgs &a, int op, hipStream_t s) {
  if (a.n <= 0 || a.N <= 0) return hipSuccess;
  if (dim == 2) return points_t<2>(a, op, s);
  if (dim == 3) return points_t<3>(a, op, s);
  return hipErrorInvalidValue;
}

}  // namespace sdot
