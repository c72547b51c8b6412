// sdot_refine.hip -- gfx950 kernels of the softmax potential transfer between
// two target levels (SoftmaxRefinement::compute_refinement with its
// local_assemble, SoftmaxRefinement.cc:61-280), FP64 throughout:
//
//   psi_f = -lam log sum_q (w_q / S_q) e^{-M_q} exp(-|x_q - y_f|^2 / (2 lam)),
//   S_q e^{M_q} = sum_c nu_c exp((psi_c - |x_q - y_c|^2 / 2) / lam),
//
// c over the coarse points within R of the center of the cell of q (the cell
// rule of the candidate query, :43-59, :81) that also pass the per-q test
// pow(|x_q - y_c|, 2) <= R*R (:152, :162, :173), f over the children of those
// coarse points that pass the same test with x_q (:99-128, :192); fine points
// that receive nothing keep psi_f = 0 (:270).
//
//   k_ref_coarse  q-major (lane = quadrature point, wave = 64 Morton-ordered
//                 points), grid.y splits the coarse bin rows into K1 segments:
//                   PASS 0: M_q = max exponent,  PASS 1: S_q against M_q,
//                   PASS 2: plain sum (use_log_sum_exp_trick off).
//   k_ref_qfin    per q, row cell order: merge the segments in fixed order,
//                 scale_q = (w_q / S_q) e^{-M_q} (0 where S_q <= 0: skipped).
//   k_ref_fine    entry-major (lane = (parent, child) entry, wave = 64
//                 entries with Morton-consecutive parents): the cells whose
//                 center is within R of the parent, their q-points within R of
//                 the child; per-segment partials, no atomics.
//   k_ref_ffin    per fine point: its entries' partials in fixed order, log.
//
// Both distance tests are exact restatements of the reference's FP64
// comparisons: s = ((0 + t0^2) + t1^2) + t2^2 without FMA contraction, then
// s <= smax (largest s with sqrt(s) <= R) for the cell rule and s <= s2max
// (largest s with pow(sqrt(s), 2) <= R*R) for the per-q test; both maps are
// monotone in s.  The exponents use s itself where the reference squares the
// rounded root (a difference of at most one rounding of s).
#include <hip/hip_runtime.h>

#include "sdot_device.hpp"

namespace sdot {

template <int DIM, int PASS>
__global__ __launch_bounds__(64) void k_ref_coarse(RefArgs a) {
  const int64_t q = (int64_t)blockIdx.x * kWave + threadIdx.x;
  const bool active = q < a.nQ;
  const int64_t qq = active ? q : a.nQ - 1;
  const int64_t cell = qq / a.nq;
  const int seg = blockIdx.y;
  double x[DIM], c[DIM], lo[DIM], hi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    x[d] = a.qxr[qq * DIM + d];
    c[d] = a.cc[cell * DIM + d];
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    lo[d] = wave_min(c[d]);
    hi[d] = wave_max(c[d]);
  }
  double M = -INFINITY, S = 0.0;
  if (PASS == 1) M = fold_partials(a.Mpart + qq, a.nQ, a.K1, -INFINITY, FoldMax{});
  cdouble *crec = cptr(a.crec);
  auto body = [&](int j0, int j1, auto) {
    for (int j = j0; j < j1; ++j) {
      cdouble *r = crec + (int64_t)j * kRec;
      double y[DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) y[d] = r[d];
      const double psi = r[3], nu = r[4];
      if (!in_set<DIM>(y, c, a.smax)) continue;
      const double s = sq_dist<DIM>(x, y);
      if (!(s <= a.s2max)) continue;
      const double t = (psi - 0.5 * s) * a.lam_inv;
      if (PASS == 0) {
        M = fmax(M, t);
      } else if (PASS == 1) {
        S += nu * exp(t - M);
      } else {
        S += nu * exp(t);
      }
    }
  };
  for_rows<DIM>(a.tg, a.tstart, lo, hi, a.Rq, 0.0, seg, a.K1, body);
  if (active) {
    if (PASS == 0)
      a.Mpart[(int64_t)seg * a.nQ + q] = M;
    else
      a.Spart[(int64_t)seg * a.nQ + q] = S;
  }
}

template <int DIM>
__global__ __launch_bounds__(256) void k_ref_qfin(RefArgs a) {
  const int64_t qr = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (qr >= a.nQ) return;
  const int64_t cr = qr / a.nq, k = qr - cr * a.nq;
  const int64_t qm = a.cell_r2m[cr] * a.nq + k;
  const double M = a.lse ? fold_partials(a.Mpart + qm, a.nQ, a.K1, -INFINITY, FoldMax{}) : -INFINITY;
  const double S = fold_partials(a.Spart + qm, a.nQ, a.K1, 0.0, FoldAdd{});
  // q-points whose sum is <= 0 contribute nothing (SoftmaxRefinement.cc:181)
  double scale = 0.0;
  if (!(S <= 0.0)) {
    scale = a.qw[qm] / S;
    if (a.lse) scale *= exp(-M);
  }
  double *r = a.rq + qr * 4;
#pragma unroll
  for (int d = 0; d < 3; ++d) r[d] = d < DIM ? a.qxr[qm * DIM + d] : 0.0;
  r[3] = scale;
}

template <int DIM>
__global__ __launch_bounds__(64) void k_ref_fine(RefArgs a) {
  // segment fastest: consecutive workgroups (one per XCD in turn) share a wave
  const int64_t blk = blockIdx.x / a.K2;
  const int seg = (int)(blockIdx.x - blk * a.K2);
  const int64_t e = blk * kWave + threadIdx.x;
  const bool active = e < a.E;
  const int64_t ee = active ? e : a.E - 1;
  double yp[DIM], yf[DIM], lo[DIM], hi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    yp[d] = a.erec[ee * kRec + d];
    yf[d] = a.erec[ee * kRec + 3 + d];
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    lo[d] = wave_min(yp[d]);
    hi[d] = wave_max(yp[d]);
  }
  double acc = 0.0;
  const int nq = a.nq;
  cdouble *rq = cptr(a.rq);
  cdouble *ccr = cptr(a.cc_r);
  auto body = [&](int c0, int c1, auto) {
    for (int cl = c0; cl < c1; ++cl) {
      double cv[DIM];
#pragma unroll
      for (int d = 0; d < DIM; ++d) cv[d] = ccr[(int64_t)cl * 4 + d];
      if (!in_set<DIM>(yp, cv, a.smax)) continue;
      cdouble *qp = rq + (int64_t)cl * nq * 4;
      for (int k = 0; k < nq; ++k) {
        double xq[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) xq[d] = qp[k * 4 + d];
        const double sc = qp[k * 4 + 3];
        const double s = sq_dist<DIM>(xq, yf);
        if (s <= a.s2max) acc += sc * exp((-0.5 * s) * a.lam_inv);
      }
    }
  };
  for_rows<DIM>(a.cg, a.cstart, lo, hi, a.Rq, 0.0, seg, a.K2, body);
  if (active) a.Epart[(int64_t)seg * a.E + e] = acc;
}

__global__ __launch_bounds__(256) void k_ref_ffin(RefArgs a) {
  const int64_t f = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (f >= a.n_fine) return;
  double pot = 0.0;
  for (int64_t i = a.foff[f]; i < a.foff[f + 1]; ++i) {
    const int64_t e = a.fent[i];
    double v = 0.0;
    for (int s = 0; s < a.K2; ++s) v += a.Epart[(int64_t)s * a.E + e];
    pot += v;
  }
  // psi_f = -lam log(pot_f) where pot_f > 0 (SoftmaxRefinement.cc:269-273)
  a.psi_fine[f] = pot > 0.0 ? -a.lam * log(pot) : pot;
}

template <int DIM>
static hipError_t refine_t(const RefArgs &a, hipStream_t s) {
  const int64_t nwq = (a.nQ + kWave - 1) / kWave;
  const dim3 gq((unsigned)nwq, (unsigned)a.K1);
  if (a.lse) {
    hipLaunchKernelGGL((k_ref_coarse<DIM, 0>), gq, dim3(64), 0, s, a);
    hipLaunchKernelGGL((k_ref_coarse<DIM, 1>), gq, dim3(64), 0, s, a);
  } else {
    hipLaunchKernelGGL((k_ref_coarse<DIM, 2>), gq, dim3(64), 0, s, a);
  }
  hipLaunchKernelGGL((k_ref_qfin<DIM>), dim3((unsigned)((a.nQ + 255) / 256)), dim3(256), 0, s, a);
  if (a.E > 0) {
    const int64_t nwe = (a.E + kWave - 1) / kWave;
    hipLaunchKernelGGL((k_ref_fine<DIM>), dim3((unsigned)(nwe * a.K2)), dim3(64), 0, s, a);
  }
  if (a.n_fine > 0)
    hipLaunchKernelGGL(k_ref_ffin, dim3((unsigned)((a.n_fine + 255) / 256)), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_refinement(int dim, const RefArgs &a, hipStream_t s) {
  if (dim == 2) return refine_t<2>(a, s);
  if (dim == 3) return refine_t<3>(a, s);
  return hipErrorInvalidValue;
}

}  // namespace sdot
