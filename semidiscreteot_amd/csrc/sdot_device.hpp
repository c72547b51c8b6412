// sdot_device.hpp -- device code shared by the single-solve kernels
// (sdot_kernels.hip) and the batched work-list kernels of the serving path
// (sdot_serve_kernels.hip): helpers, argument blocks and the bodies of the
// evaluation passes.  A body receives its block coordinates as arguments, so the
// same code runs from a one-solve grid and from a multi-solve work list.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

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

// Squared distance ((0 + t0^2) + t1^2) + t2^2 without FMA contraction: the
// reference's FP64 sum of squares under its square root.
template <int DIM>
__device__ __forceinline__ double sq_dist(const double *a, const double *b) {
#pragma clang fp contract(off)
  double s = 0.0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const double t = a[d] - b[d];
    s = s + t * t;
  }
  return s;
}

// The value of lane 0 in every lane, held in scalar registers.
__device__ __forceinline__ double wave_uniform(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(unsigned)(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmin(v, __shfl_xor(v, m));
  return wave_uniform(v);
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m));
  return wave_uniform(v);
}

// Records a kernel only reads are addressed through the constant address
// space: the compiler may then keep wave-uniform record loads in scalar
// registers (s_load) even where a store of the same kernel could alias them
// as far as it can tell (the persistent work-list loops store partial sums
// between the record streams of consecutive items).
using cdouble = const __attribute__((address_space(4))) double;

__device__ __forceinline__ cdouble *cptr(const double *p) { return (cdouble *)p; }

template <int MODE>
__device__ __forceinline__ double expo(double d) {
  if constexpr (MODE == kF32Lse) {
    return (double)__builtin_amdgcn_exp2f((float)d);
  } else {
    return exp(d);
  }
}

// v_max_f64 / v_min_f64 as one instruction: fmax() on a loop-carried value
// also canonicalizes that operand (a second v_max_f64) to quiet signaling
// NaNs, which the exponents and maxima here never are.
__device__ __forceinline__ double vmax64(double a, double b) {
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ double vmin64(double a, double b) {
  double r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

// Records per chunk of the exponent-change extrema (seed bound of the FP32 sweep).
constexpr int kChunk = 64;

using Tested = std::true_type;    // run whose members need the candidate test
using Inside = std::false_type;   // run of candidates of every query point

// Visit the bin rows (x-runs of bins, contiguous in row-major bin order) whose
// bins can hold a point within Rq of the box [lo,hi]; the rows whose global
// index (bz * n_y + by) is seg modulo nseg are handed to body(first_item,
// end_item, tag).  The split depends on the row alone, not on the box, so the
// segments a workgroup of one XCD runs (seg = xcd modulo 8) touch a fixed
// eighth of the rows whatever wave they serve, and that eighth of the records
// stays in the XCD's L2.  With s_in > 0 a row
// is split into runs: bins every point of which lies within sqrt(s_in) of every
// point of the box (tag Inside: candidates of every query point in the box,
// s_in being the candidate bound with a relative margin far above the rounding
// of the squared distance) and the rest (tag Tested).
template <int DIM, class Body>
__device__ __forceinline__ void for_rows(const Grid &g, const int32_t *__restrict__ start,
                                         const double *lo, const double *hi, double Rq,
                                         double s_in, int seg, int nseg, Body &&body) {
  int b0[3] = {0, 0, 0}, b1[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    b0[d] = bin_of(lo[d] - Rq, g.o[d], g.inv_h, g.n[d]);
    b1[d] = bin_of(hi[d] + Rq, g.o[d], g.inv_h, g.n[d]);
  }
  const double R2 = Rq * Rq;
  const double pad = 1e-9 * g.h;
  for (int bz = b0[2]; bz <= b1[2]; ++bz) {
    double dz = 0.0, dzf = 0.0;
    if (DIM == 3) {
      const double zl = g.o[2] + bz * g.h, zh = zl + g.h;
      dz = fmax(0.0, fmax(zl - hi[2], lo[2] - zh));
      dzf = fmax(zh - lo[2], hi[2] - zl);
    }
    // this segment's rows of the slab: by = seg - bz * n_y (mod nseg)
    const int want = ((seg - (bz * g.n[1]) % nseg) % nseg + nseg) % nseg;
    const int by0 = b0[1] + ((want - b0[1] % nseg) % nseg + nseg) % nseg;
    for (int by = by0; by <= b1[1]; by += nseg) {
      const double yl = g.o[1] + by * g.h, yh = yl + g.h;
      const double dy = fmax(0.0, fmax(yl - hi[1], lo[1] - yh));
      const double r2 = R2 - dy * dy - dz * dz;
      if (r2 < 0.0) continue;
      const double wx = sqrt(r2);
      const int x0 = bin_of(lo[0] - wx, g.o[0], g.inv_h, g.n[0]);
      const int x1 = bin_of(hi[0] + wx, g.o[0], g.inv_h, g.n[0]);
      const int base = (bz * g.n[1] + by) * g.n[0];
      // inner bins [i0, i1]: farthest y / z distance to the box, then the x
      // interval [hi - w, lo + w] every point of which is within sqrt(s_in) of
      // every point of the box; only bins entirely inside it (less a pad that
      // covers the rounding of the binning rule)
      int i0 = x1 + 1, i1 = x1;
      if (s_in > 0.0) {
        const double dyf = fmax(yh - lo[1], hi[1] - yl);
        const double r2x = s_in - dyf * dyf - dzf * dzf;
        if (r2x > 0.0) {
          const double w = sqrt(r2x);
          const double xa = hi[0] - w + pad, xb = lo[0] + w - pad;
          if (xa < xb) {
            i0 = max(x0, (int)ceil((xa - g.o[0]) * g.inv_h));
            i1 = min(x1, (int)floor((xb - g.o[0]) * g.inv_h) - 1);
          }
        }
      }
      const int s = __builtin_amdgcn_readfirstlane(start[base + x0]);
      const int e = __builtin_amdgcn_readfirstlane(start[base + x1 + 1]);
      if (i0 <= i1) {
        const int m0 = __builtin_amdgcn_readfirstlane(start[base + i0]);
        const int m1 = __builtin_amdgcn_readfirstlane(start[base + i1 + 1]);
        if (s < m0) body(s, m0, Tested{});
        if (m0 < m1) body(m0, m1, Inside{});
        if (m1 < e) body(m1, e, Tested{});
      } else if (s < e) {
        body(s, e, Tested{});
      }
    }
  }
}

// Value bound of a traversal.  Every pair (q, j) a sweep sums satisfies
//   psi~_j - |x_q - y_j|^2 / 2  >=  V_q
// or its term is below 2^-theta of the terms the sum keeps (psi~_j = psi_j +
// eps ln nu_j; V_q from the q-point's largest exponent, see pass1_body and
// grad_body).  For a wave with q-point box B (q-major) or target box B
// (target-major), a bin b of the traversed grid can hold such a pair only if
//   A_b - d(B, bin b)^2 / 2  >=  B_b
// with A the largest psi~ and B the smallest V on either side: per bin and per
// row of the grid from arrays, or one value for the whole wave.  Bins that fail
// are skipped whole.
struct VBound {
  const double *abin = nullptr, *arow = nullptr;  // per bin / per row A (null: auni)
  const double *bbin = nullptr, *brow = nullptr;  // per bin / per row B (null: buni)
  double auni = 0.0, buni = 0.0;
  double lo[3] = {0.0, 0.0, 0.0}, hi[3] = {0.0, 0.0, 0.0};  // the wave's box, raw coordinates
};

// Gap between the intervals [lo, hi] and [a, b] (0 when they overlap).
__device__ __forceinline__ double gap(double lo, double hi, double a, double b) {
  return fmax(0.0, fmax(a - hi, lo - b));
}

// for_rows restricted by a value bound: the same rows and segments (row index
// modulo nseg), the same candidate enumeration radius Rq and inner runs, and
// only the bins vb admits; runs of consecutive admitted bins go to body.
// dense: every target is a candidate (all bins, no candidate test).
template <int DIM, class Body>
__device__ __forceinline__ void for_rows_vb(const Grid &g, const int32_t *__restrict__ start,
                                            const double *lo, const double *hi, double Rq,
                                            double s_in, int seg, int nseg, bool dense,
                                            const VBound &vb, Body &&body) {
  cdouble *abin = cptr(vb.abin), *arow = cptr(vb.arow);
  cdouble *bbin = cptr(vb.bbin), *brow = cptr(vb.brow);
  int b0[3] = {0, 0, 0}, b1[3] = {0, 0, 0};
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    b0[d] = dense ? 0 : bin_of(lo[d] - Rq, g.o[d], g.inv_h, g.n[d]);
    b1[d] = dense ? g.n[d] - 1 : bin_of(hi[d] + Rq, g.o[d], g.inv_h, g.n[d]);
  }
  const double R2 = Rq * Rq;
  const double pad = 1e-9 * g.h;
  for (int bz = b0[2]; bz <= b1[2]; ++bz) {
    double dz = 0.0, dzf = 0.0, vdz = 0.0;
    if (DIM == 3) {
      const double zl = g.o[2] + bz * g.h, zh = zl + g.h;
      dz = fmax(0.0, fmax(zl - hi[2], lo[2] - zh));
      dzf = fmax(zh - lo[2], hi[2] - zl);
      vdz = gap(vb.lo[2], vb.hi[2], zl, zh);
    }
    const int want = ((seg - (bz * g.n[1]) % nseg) % nseg + nseg) % nseg;
    const int by0 = b0[1] + ((want - b0[1] % nseg) % nseg + nseg) % nseg;
    for (int by = by0; by <= b1[1]; by += nseg) {
      const double yl = g.o[1] + by * g.h, yh = yl + g.h;
      const int row = bz * g.n[1] + by;
      // value bound of the row: the x-interval of bins it can admit
      const double vd2 = wave_uniform(gap(vb.lo[1], vb.hi[1], yl, yh) *
                                      gap(vb.lo[1], vb.hi[1], yl, yh) + vdz * vdz);
      const double ar = arow ? arow[row] : vb.auni;
      const double br = brow ? brow[row] : vb.buni;
      const double slack = ar - br - 0.5 * vd2;
      if (!(slack >= 0.0)) continue;
      const double wv = sqrt(2.0 * slack);
      int x0 = bin_of(vb.lo[0] - wv, g.o[0], g.inv_h, g.n[0]);
      int x1 = bin_of(vb.hi[0] + wv, g.o[0], g.inv_h, g.n[0]);
      int i0 = x1 + 1, i1 = x1;
      if (dense) {
        i0 = x0;  // every bin is inside
      } else {
        const double dy = fmax(0.0, fmax(yl - hi[1], lo[1] - yh));
        const double r2 = R2 - dy * dy - dz * dz;
        if (r2 < 0.0) continue;
        const double wx = sqrt(r2);
        x0 = max(x0, bin_of(lo[0] - wx, g.o[0], g.inv_h, g.n[0]));
        x1 = min(x1, bin_of(hi[0] + wx, g.o[0], g.inv_h, g.n[0]));
        i0 = x1 + 1;
        i1 = x1;
        if (s_in > 0.0) {
          const double dyf = fmax(yh - lo[1], hi[1] - yl);
          const double r2x = s_in - dyf * dyf - dzf * dzf;
          if (r2x > 0.0) {
            const double w = sqrt(r2x);
            const double xa = hi[0] - w + pad, xb = lo[0] + w - pad;
            if (xa < xb) {
              i0 = max(x0, (int)ceil((xa - g.o[0]) * g.inv_h));
              i1 = min(x1, (int)floor((xb - g.o[0]) * g.inv_h) - 1);
            }
          }
        }
      }
      if (x0 > x1) continue;
      const int base = row * g.n[0];
      // runs of consecutive admitted bins, split into tested and inside parts
      auto emit = [&](int r0, int r1) {
        const int s = __builtin_amdgcn_readfirstlane(start[base + r0]);
        const int e = __builtin_amdgcn_readfirstlane(start[base + r1 + 1]);
        if (s >= e) return;
        const int j0 = max(r0, i0), j1 = min(r1, i1);
        if (j0 <= j1) {
          const int m0 = __builtin_amdgcn_readfirstlane(start[base + j0]);
          const int m1 = __builtin_amdgcn_readfirstlane(start[base + j1 + 1]);
          if (s < m0) body(s, m0, Tested{});
          if (m0 < m1) body(m0, m1, Inside{});
          if (m1 < e) body(m1, e, Tested{});
        } else {
          body(s, e, Tested{});
        }
      };
      int rs = -1;
      for (int b = x0; b <= x1; ++b) {
        const double xl = g.o[0] + b * g.h;
        const double dx = gap(vb.lo[0], vb.hi[0], xl, xl + g.h);
        const double ab = abin ? abin[base + b] : vb.auni;
        const double bb = bbin ? bbin[base + b] : vb.buni;
        const bool keep = ab - bb - 0.5 * (dx * dx + vd2) >= 0.0;
        if (keep && rs < 0) rs = b;
        if (!keep && rs >= 0) {
          emit(rs, b - 1);
          rs = -1;
        }
      }
      if (rs >= 0) emit(rs, x1);
    }
  }
}

// Fold f over the K segment partials p[k * stride], k = 0 .. K-1, in index
// order.  The loads of 16 partials are issued together: a plain loop waits
// out one memory round trip per partial, which dominates the finishing
// kernels of small grids.  Same operations in the same order as the plain
// loop, so the result is bitwise the same.
template <class F>
__device__ __forceinline__ double fold_partials(const double *__restrict__ p, int64_t stride,
                                                int K, double init, F &&f) {
  constexpr int B = 16;
  double r = init;
  for (int k0 = 0; k0 < K; k0 += B) {
    double v[B];
#pragma unroll
    for (int i = 0; i < B; ++i) v[i] = (k0 + i < K) ? p[(int64_t)(k0 + i) * stride] : 0.0;
#pragma unroll
    for (int i = 0; i < B; ++i)
      if (k0 + i < K) r = f(r, v[i]);
  }
  return r;
}
struct FoldAdd {
  __device__ __forceinline__ double operator()(double a, double b) const { return a + b; }
};
struct FoldMax {
  __device__ __forceinline__ double operator()(double a, double b) const { return fmax(a, b); }
};

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

// FP32 sweep (PASS 2).  Every exponent t of a work item (a block of
// quadrature points x a segment of candidate targets) is taken relative to an
// integer reference value M that bounds the item's exponents from above, so
// that the dominant terms have t - M in [-(spread + 1), 0] and the FP32
// argument keeps its precision (the rounding of t - M to FP32 costs
// |t - M| 2^-25 ln 2 of relative error per term).  M is ceil of either the
// item's exact maximum (a max loop over its targets first) or, when the
// previous evaluation of the same item is available, of its previous maximum
// raised by the largest exponent change among its targets: an upper bound of
// the new maximum, at most (spread of the changes) + 1 above it.  The seed is
// used while that spread, over the item's own targets, stays below
// kSeedSpread log2 units.  When the radius changed since that evaluation, the
// item's candidates may have changed too: a target that joined can lie above
// the seed, and an item whose maximum exceeds it by more than kSeedAbove (or
// fell further than the spread allows below it) is swept again against its
// exact maximum.
constexpr double kSeedSpread = 2.0;
constexpr double kSeedAbove = 0.5;

// Target record (64 bytes):
//   [0]=a  [1..DIM]=y'/eps  [1+DIM..2DIM]=y raw  [1+2DIM]=nu
template <int DIM>
struct TRec {
  double a, yl[DIM], y[DIM], nu;
  template <class P>
  __device__ __forceinline__ void load(P r) {
    a = r[0];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      yl[d] = r[1 + d];
      y[d] = r[1 + DIM + d];
    }
    nu = r[1 + 2 * DIM];
  }
  // the whole record in one 64-byte scalar load, issued ahead of its use: the
  // field loads of load() are placed at their uses, which puts one scalar
  // round trip before the candidate test and another after it
  __device__ __forceinline__ void load_whole(cdouble *r) {
    typedef double d8 __attribute__((ext_vector_type(8)));
    const d8 v = *reinterpret_cast<const __attribute__((address_space(4))) d8 *>(r);
    a = v[0];
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      yl[d] = v[1 + d];
      y[d] = v[1 + DIM + d];
    }
    nu = v[1 + 2 * DIM];
  }
  // the fields of the exponent and the weight only (a, yl, nu: y is left
  // unset), each load of them wholly used: a record load with unused parts
  // lets the next record's address arithmetic reuse those registers, which
  // makes it wait for the load
  __device__ __forceinline__ void load_core(cdouble *r) {
    if constexpr (DIM == 3) {
      typedef double d4 __attribute__((ext_vector_type(4)));
      const d4 v = *reinterpret_cast<const __attribute__((address_space(4))) d4 *>(r);
      a = v[0];
      yl[0] = v[1];
      yl[1] = v[2];
      yl[2] = v[3];
    } else {
      typedef double d2 __attribute__((ext_vector_type(2)));
      const d2 v = *reinterpret_cast<const __attribute__((address_space(4))) d2 *>(r);
      a = v[0];
      yl[0] = v[1];
      yl[1] = r[2];
    }
    nu = r[1 + 2 * DIM];
  }
};

// Records streamed two at a time: both records' scalar loads are issued
// together and waited for once, which halves the load round trips a wave
// waits out per record (a load issued one record ahead is sunk by the
// compiler to its use, and scalar loads complete out of order, so any wait
// covers every load in flight).  The lanes of a record: every lane
// (TESTED false) or the lanes whose point set c holds it (the candidate test,
// for both records before either's arithmetic, so that no load is sunk below
// it).
template <int DIM, bool TESTED, class Body>
__device__ __forceinline__ void stream_records(cdouble *trec, int j0, int j1, const double *c,
                                               double smax, Body &&body) {
  auto load = [&](TRec<DIM> &r, int j) {
    if constexpr (TESTED)
      r.load_whole(trec + (int64_t)j * kRec);
    else
      r.load_core(trec + (int64_t)j * kRec);
  };
  int j = j0;
  for (; j + 1 < j1; j += 2) {
    TRec<DIM> A, B;
    load(A, j);
    load(B, j + 1);
    // both loads before the first use (the scheduler would delay the second
    // past the wait for the first)
    __builtin_amdgcn_sched_barrier(0);
    if constexpr (TESTED) {
      const uint64_t ma = __ballot(in_set<DIM>(A.y, c, smax));
      const uint64_t mb = __ballot(in_set<DIM>(B.y, c, smax));
      if (__builtin_amdgcn_inverse_ballot_w64(ma)) body(A);
      if (__builtin_amdgcn_inverse_ballot_w64(mb)) body(B);
    } else {
      body(A);
      body(B);
    }
  }
  if (j < j1) {
    TRec<DIM> A;
    load(A, j);
    if (!TESTED || in_set<DIM>(A.y, c, smax)) body(A);
  }
}

// ---------------------------------------------------------------------------
// k_prep: target records
// ---------------------------------------------------------------------------

template <int DIM>
__device__ __forceinline__ void prep_body(const PrepArgs &a, int64_t blk, double *sh) {
  const int64_t j = blk * 256 + threadIdx.x;
  double pn = 0.0;
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
    // value bound: psi + eps ln nu per target, row order
    if (a.psit_r) a.psit_r[a.pos_r[j]] = psi + a.eps * log(nu);
    if (a.psi_prev) {
      // change of the exponent coefficient since the previous evaluation, in
      // the row order the q-major sweep streams
      if (a.hist) a.dpsi_r[a.pos_r[j]] = (psi - a.psi_prev[j]) * a.coef;
      a.psi_prev[j] = psi;
    }
  }
  const double s = block_sum_256(pn, sh);
  if (threadIdx.x == 0) a.PNpart[blk] = s;
}

// Extrema of the exponent change dpsi_r over each chunk of kChunk row-order
// records (one wave per chunk).
__device__ __forceinline__ void chunkext_body(const double *dpsi_r, int64_t N, double *dcx,
                                              double *dcn, int64_t chunk) {
  const int64_t j = chunk * kChunk + (threadIdx.x & (kWave - 1));
  double mx = -INFINITY, mn = INFINITY;
  if (j < N) mx = mn = dpsi_r[j];
#pragma unroll
  for (int m = kWave / 2; m >= 1; m >>= 1) {
    mx = fmax(mx, __shfl_xor(mx, m));
    mn = fmin(mn, __shfl_xor(mn, m));
  }
  if ((threadIdx.x & (kWave - 1)) == 0 && chunk * kChunk < N) {
    dcx[chunk] = mx;
    dcn[chunk] = mn;
  }
}

// Reductions of the value bound (see VBound).  Over the records / cells of bin
// b (offsets start[b] .. start[b+1]): the largest psi~ of a target bin (empty:
// -inf) or the smallest V of a cell bin (empty: +inf).
// One wave per bin (clustered targets put hundreds of records in some bins):
// the lanes stride over the bin's entries, then a wave reduction.
template <bool MAX>
__device__ __forceinline__ void bin_reduce_body(const double *v, const int32_t *start, int64_t nb,
                                                double *out, int64_t b) {
  const int lane = threadIdx.x & (kWave - 1);
  double r = MAX ? -INFINITY : INFINITY;
  if (b < nb)
    for (int32_t i = start[b] + lane; i < start[b + 1]; i += kWave)
      r = MAX ? fmax(r, v[i]) : fmin(r, v[i]);
  r = MAX ? wave_max(r) : wave_min(r);
  if (b < nb && lane == 0) out[b] = r;
}

// ... and over the nx bins of each row of the grid.
template <bool MAX>
__device__ __forceinline__ void row_reduce_body(const double *bins, int64_t nrows, int nx,
                                                double *out, int64_t row) {
  if (row >= nrows) return;
  double r = MAX ? -INFINITY : INFINITY;
  for (int i = 0; i < nx; ++i) {
    const double v = bins[row * nx + i];
    r = MAX ? fmax(r, v) : fmin(r, v);
  }
  out[row] = r;
}

// Smallest V_q = (M_q + log S_q - theta)/c - |x'_q|^2/2 over the quadrature
// points of each cell (row cell order), from the q records; q-points that
// received nothing need nothing (+inf).  log2 S in the FP32 sweep's units.
template <int DIM>
__device__ __forceinline__ void cellv_body(const double *qrec_r, int64_t n_cells, int nq,
                                           double theta, double inv_coef, bool log2_units,
                                           double *vcell, int64_t c) {
  if (c >= n_cells) return;
  double v = INFINITY;
  for (int k = 0; k < nq; ++k) {
    const double *r = qrec_r + (c * nq + k) * kRec;
    const double M = r[DIM], S = r[DIM + 2];
    if (!(S > 0.0) || !(M > -INFINITY) || !(M < INFINITY)) continue;
    double x2 = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) x2 = fma(r[d], r[d], x2);
    const double ls = log2_units ? log2(S) : log(S);
    v = fmin(v, (M + ls - theta) * inv_coef - 0.5 * x2);
  }
  // rounding of the bound: a relative margin far below 2^-theta
  vcell[c] = v - 1e-12 * (1.0 + fabs(v));
}

// ---------------------------------------------------------------------------
// k_pass1: q-major sweeps
// ---------------------------------------------------------------------------

template <int DIM, int G, int MODE, int PASS, bool VB = false>
__device__ __forceinline__ void pass1_body(const P1Args &a, int64_t blk, int seg) {
  const int lane = threadIdx.x % kWave;
  const int64_t grp = blk * kWave + lane;
  const bool active = grp < a.n_groups;
  const int64_t gg = active ? grp : 0;
  const int64_t cell = gg / a.gpc;
  const int64_t q0 = cell * a.nq + (gg - cell * a.gpc) * G;

  double c[DIM], x[G][DIM], M[G], S[G], T[G];
#pragma unroll
  for (int d = 0; d < DIM; ++d) c[d] = a.cc[cell * DIM + d];
#pragma unroll
  for (int k = 0; k < G; ++k) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) x[k][d] = a.qx[(q0 + k) * DIM + d];
    S[k] = 0.0;
    T[k] = -INFINITY;
    if (PASS == 0 || PASS == 2) {
      M[k] = -INFINITY;
    } else if (MODE == kF64Plain) {
      M[k] = a.plain_coef * a.qx2[q0 + k];
    } else {
      double m = -INFINITY;
      if (a.Mglob)
        m = a.Mglob[q0 + k];
      else
        for (int s = 0; s < a.K1; ++s) m = fmax(m, a.Mpart[(int64_t)s * a.nQ + q0 + k]);
      M[k] = m;
    }
  }
  double lo[DIM], hi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    lo[d] = a.blkbox[blk * 2 * DIM + d];
    hi[d] = a.blkbox[blk * 2 * DIM + DIM + d];
  }
  // one candidate record on the lanes that hold it (records() applies the test)
  auto one = [&](const TRec<DIM> &r) {
    if (PASS == 0) {
#pragma unroll
      for (int k = 0; k < G; ++k) {
        double t = r.a;
#pragma unroll
        for (int d = 0; d < DIM; ++d) t = fma(x[k][d], r.yl[d], t);
        M[k] = vmax64(M[k], t);
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
      // exponent relative to the bound M (<= 0 up to rounding), running
      // maximum; the terms are accumulated in FP64
#pragma unroll
      for (int k = 0; k < G; ++k) {
        double t = r.a;
#pragma unroll
        for (int d = 0; d < DIM; ++d) t = fma(x[k][d], r.yl[d], t);
        T[k] = vmax64(T[k], t);
        S[k] = fma(r.nu, (double)__builtin_amdgcn_exp2f((float)(t - M[k])), S[k]);
      }
    }
  };
  auto sweep = [&](auto &&f) {
    if (VB && a.dense) {
      // the segments of the bounded sums are sets of bin rows: the item's
      // maxima and seed extrema must cover the same targets
      VBound open;
      open.auni = INFINITY;
      open.buni = -INFINITY;
      for_rows_vb<DIM>(a.tg, a.tstart, lo, hi, a.Rq, a.s_in, seg, a.K1, true, open, f);
    } else if (a.dense) {
      const int64_t j0 = a.N * seg / a.K1, j1 = a.N * (seg + 1) / a.K1;
      f((int)j0, (int)j1, Inside{});
    } else {
      for_rows<DIM>(a.tg, a.tstart, lo, hi, a.Rq, a.s_in, seg, a.K1, f);
    }
  };
  // Value bound of the sums (VB): with L_q a lower bound of the q-point's
  // largest exponent (in the sweep's units, c = a.coef per unit of psi), a pair
  // whose weighted term nu_j e^{t_qj} lies below 2^-theta nu_min e^{L_q} is
  // skipped; this is psi~_j - |x_q - y_j|^2/2 < V_q with
  //   V_q = (L_q - theta) / c + eps ln nu_min - |x'_q|^2 / 2.
  // Every skipped term is then below 2^-theta of the q-point's sum (which holds
  // its largest term, nu >= nu_min), so the sum loses less than N 2^-theta of
  // itself and a target's mass less than 2^-theta.  vb takes the smallest V_q
  // of the wave and the box of its points; the bins' largest psi~ are a.pbin /
  // a.prow.
  VBound vb;
  auto set_bound = [&](const double (&L)[G], double theta) {
    double v = INFINITY;
#pragma unroll
    for (int k = 0; k < G; ++k) {
      double x2 = 0.0;
#pragma unroll
      for (int d = 0; d < DIM; ++d) x2 = fma(x[k][d], x[k][d], x2);
      // a q-point without candidates in this segment needs nothing from it
      const double vk = (L[k] > -INFINITY)
                            ? (L[k] - theta) * a.inv_coef + a.vshift - 0.5 * x2
                            : INFINITY;
      v = fmin(v, vk);
    }
    if (!active) v = INFINITY;
    // rounding of the bound: a relative margin far below 2^-theta; the wave's
    // values are kept in scalar registers (the sweep's VGPRs bound its occupancy)
    v = wave_min(v);
    vb.buni = wave_uniform(v - 1e-12 * (1.0 + fabs(v)));
    // the box of the block's q-points (host-computed, raw coordinates)
    cdouble *qb = cptr(a.qbox) + blk * 2 * DIM;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      vb.lo[d] = qb[d];
      vb.hi[d] = qb[DIM + d];
    }
    vb.abin = a.pbin;
    vb.arow = a.prow;
  };
  // the maxima of PASS 0 and the plain mode's sums are not bounded
  constexpr bool vbound = VB && PASS != 0 && MODE != kF64Plain;
  auto sweep_v = [&](auto &&f) {
    if constexpr (vbound)
      for_rows_vb<DIM>(a.tg, a.tstart, lo, hi, a.Rq, a.s_in, seg, a.K1, a.dense != 0, vb, f);
    else
      sweep(f);
  };
  // the item's candidate records, with scalar loads
  cdouble *trec = cptr(a.trec);
  auto records = [&](auto &&fn, bool bounded) {
    auto runs = [&](int j0, int j1, auto tag) {
      stream_records<DIM, decltype(tag)::value>(trec, j0, j1, c, a.smax, fn);
    };
    if (bounded)
      sweep_v(runs);
    else
      sweep(runs);
  };
  // exact maximum over the item's candidates (every candidate: no value bound)
  // (bounded: only the bins the current bound admits)
  auto max_loop = [&](bool bounded) {
    records(
        [&](const TRec<DIM> &r) {
#pragma unroll
          for (int k = 0; k < G; ++k) {
            double t = r.a;
#pragma unroll
            for (int d = 0; d < DIM; ++d) t = fma(x[k][d], r.yl[d], t);
            T[k] = vmax64(T[k], t);
          }
        },
        bounded);
  };

  double L[G];  // lower bounds of the largest exponents (value bound)
#pragma unroll
  for (int k = 0; k < G; ++k) L[k] = M[k];
  bool seedq = false;  // the item's reference value came from the history seed
  if (PASS == 2) {
    // reference value of this item (see kSeedSpread)
    bool seed = false;
    double dmx = -INFINITY, dmn = INFINITY;
    if (a.hist) {
      if (a.dcx) {
        // extrema per chunk of kChunk records: a bound over a superset of
        // the item's candidates
        cdouble *cx = cptr(a.dcx);
        cdouble *cn = cptr(a.dcn);
        sweep([&](int j0, int j1, auto) {
          if (j1 <= j0) return;
          for (int ch = j0 / kChunk; ch <= (j1 - 1) / kChunk; ++ch) {
            dmx = vmax64(dmx, cx[ch]);
            dmn = vmin64(dmn, cn[ch]);
          }
        });
      } else {
        cdouble *dp = cptr(a.dpsi);
        sweep([&](int j0, int j1, auto) {
          for (int j = j0; j < j1; ++j) {
            dmx = vmax64(dmx, dp[j]);
            dmn = vmin64(dmn, dp[j]);
          }
        });
      }
      seed = dmx - dmn < kSeedSpread;
    }
    // the previous maxima bound the new ones from below (value bound) when the
    // item's candidates are the same or more: at a larger radius, unless a
    // point had no candidate in the item before (no bound at all)
    bool hist_lb = a.hist != 0;
    if constexpr (vbound) {
      if (a.hist && !a.hist_same) {
        bool empty = false;
#pragma unroll
        for (int k = 0; k < G; ++k)
          empty |= active && !(a.Tout[(int64_t)seg * a.nQ + q0 + k] > -INFINITY);
        hist_lb = __ballot(empty) == 0;
      }
      seed = seed && hist_lb;
    }
    seedq = seed;
    if (seed) {
      // the new maximum lies in [T_prev + dmn, T_prev + dmx]
#pragma unroll
      for (int k = 0; k < G; ++k) {
        const double tp = a.Tout[(int64_t)seg * a.nQ + q0 + k];
        M[k] = ceil(tp + dmx) + a.seed_bias;
        L[k] = tp + dmn;
      }
    } else {
      bool bounded = false;
      if constexpr (vbound) {
        if (hist_lb) {
          // T_prev + dmn is a lower bound of the new maximum: a bin whose
          // exponents all lie below it cannot hold the maximum (the bound at
          // 0 bits, the weights' spread included)
#pragma unroll
          for (int k = 0; k < G; ++k) L[k] = a.Tout[(int64_t)seg * a.nQ + q0 + k] + dmn;
          set_bound(L, 0.0);
          bounded = true;
        }
      }
      max_loop(bounded);
#pragma unroll
      for (int k = 0; k < G; ++k) {
        M[k] = ceil(T[k]);
        L[k] = T[k];
        T[k] = -INFINITY;
      }
    }
  }
  if constexpr (vbound) set_bound(L, a.theta);
  records(one, true);
  if (PASS == 2 && seedq) {
    // With the same candidates the seeded bound holds by construction: the
    // new maximum lies at most M and less than kSeedSpread + 1 below it.  An
    // item whose maximum fell further (history that does not belong to these
    // potentials) or rose above M (a target that joined at a larger radius) is
    // swept again against its exact maximum, so that its dominant terms keep
    // their FP32 argument precision; the maximum comes from a sweep over every
    // candidate, since the value bound of the first sweep rested on the history.
    bool low = false;
#pragma unroll
    for (int k = 0; k < G; ++k)
      low |= T[k] < M[k] - (kSeedSpread + 2.0) || T[k] > M[k] + kSeedAbove;
    if (__ballot(low) != 0) {
#pragma unroll
      for (int k = 0; k < G; ++k) T[k] = -INFINITY;
      max_loop(false);
#pragma unroll
      for (int k = 0; k < G; ++k) {
        M[k] = ceil(T[k]);
        L[k] = T[k];
        S[k] = 0.0;
        T[k] = -INFINITY;
      }
      if constexpr (vbound) set_bound(L, a.theta);
      records(one, true);
    }
  }
  if (active) {
#pragma unroll
    for (int k = 0; k < G; ++k) {
      const int64_t o = (int64_t)seg * a.nQ + q0 + k;
      if (PASS == 0) a.Mout[o] = M[k];
      if (PASS == 1) a.Sout[o] = S[k];
      if (PASS == 2) {
        a.Mout[o] = M[k];
        a.Sout[o] = S[k];
        a.Tout[o] = T[k];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// k_qfin: per-q merge, functional, scale, q records (row cell order)
// ---------------------------------------------------------------------------

template <int DIM, int MODE>
__device__ __forceinline__ void qfin_body(const QFArgs &a, int64_t blk, double *sh) {
  const int64_t qr = blk * 256 + threadIdx.x;
  double Fq = 0.0, Cq = 0.0, Uq = 0.0;
  if (qr < a.nQ) {
    const int64_t cr = qr / a.nq, k = qr - cr * a.nq;
    const int64_t qm = a.cell_r2m[cr] * a.nq + k;
    double x[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) x[d] = a.qx[qm * DIM + d];
    const double x2 = a.qx2[qm];
    double M, S = 0.0, T = 0.0;
    if (MODE == kF64Plain) {
      M = a.plain_coef * x2;
      S = fold_partials(a.Spart + qm, a.nQ, a.K1, 0.0, FoldAdd{});
    } else if (MODE == kF64Lse) {
      M = fold_partials(a.Mpart + qm, a.nQ, a.K1, -INFINITY, FoldMax{});
      S = fold_partials(a.Spart + qm, a.nQ, a.K1, 0.0, FoldAdd{});
    } else {
      // segments carry integer reference values: S = sum_s S_s 2^(M_s - M),
      // then renormalised to M = ceil(largest exponent) for the gradient pass
      M = fold_partials(a.Mpart + qm, a.nQ, a.K1, -INFINITY, FoldMax{});
      T = fold_partials(a.Tpart + qm, a.nQ, a.K1, -INFINITY, FoldMax{});
      if (M > -INFINITY && T > -INFINITY) {
        // segment order, with the loads of 8 segments issued together
        constexpr int B = 8;
        for (int s0 = 0; s0 < a.K1; s0 += B) {
          double sv[B], mv[B];
#pragma unroll
          for (int i = 0; i < B; ++i) {
            const bool in = s0 + i < a.K1;
            sv[i] = in ? a.Spart[(int64_t)(s0 + i) * a.nQ + qm] : 0.0;
            mv[i] = in ? a.Mpart[(int64_t)(s0 + i) * a.nQ + qm] : 0.0;
          }
#pragma unroll
          for (int i = 0; i < B; ++i)
            if (sv[i] > 0.0) S += ldexp(sv[i], (int)(mv[i] - M));
        }
        const double Mt = ceil(T);
        S = ldexp(S, (int)(M - Mt));
        M = Mt;
      } else {
        M = -INFINITY;  // no candidates
      }
      // safety net: every term far below the reference value (cannot happen
      // while the seed bound holds); the host repeats the evaluation unseeded
      if (M > -INFINITY && S < 1e-40) Uq = 1.0;
    }
    const double w = a.qw[qm];
    double Mrec = INFINITY, scale = 0.0;
    // q-points whose sum vanished are skipped (SotSolver.cc:854)
    if (S > 0.0 && M > -INFINITY) {
      Fq = w * (a.eps * (M * a.ln_scale + log(S)) - 0.5 * x2);
      scale = w / S;
      // C_global sums w_q / S_q with S_q taken relative to the largest term
      // (SotSolver.cc:290); the FP32 sum is relative to 2^M = 2^ceil(T) instead
      Cq = MODE == kF32Lse ? scale * exp2(T - M) : scale;
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
    if (a.qchunk) {
      const int64_t ch = qr / a.G, k = qr - ch * a.G;
      double *cq = a.qchunk + ch * (DIM + 2) * a.G + k;
#pragma unroll
      for (int d = 0; d < DIM; ++d) cq[d * a.G] = x[d];
      cq[DIM * a.G] = Mrec;
      cq[(DIM + 1) * a.G] = scale;
    }
  }
  const double fs = block_sum_256(Fq, sh);
  const double cs = block_sum_256(Cq, sh);
  const double us = block_sum_256(Uq, sh);
  if (threadIdx.x == 0) {
    a.Fpart[blk] = fs;
    a.Cpart[blk] = cs;
    a.Upart[blk] = us;
  }
}

// ---------------------------------------------------------------------------
// k_grad: target-major gradient partials
// ---------------------------------------------------------------------------

template <int DIM, int MODE, int G, bool BARY = false, bool VB = false>
__device__ __forceinline__ void grad_body(const GArgs &a, int64_t blk, int seg) {
  const int64_t j = blk * kWave + threadIdx.x;
  const bool active = j < a.N;
  TRec<DIM> me;
  me.load(a.trec_m + (active ? j : 0) * kRec);
  double lo[DIM], hi[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    lo[d] = a.twbox[blk * 2 * DIM + d];
    hi[d] = a.twbox[blk * 2 * DIM + DIM + d];
  }
  double acc = 0.0;
  double accx[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) accx[d] = 0.0;
  const int nq = a.nq;
  cdouble *qchk = cptr(a.qchunk);
  cdouble *ccr = cptr(a.cc_r);
  // one chunk of G q-points (GArgs::qchunk layout: x'_d of the G points, then
  // their M, then their w/S)
  auto chunk = [&](cdouble *qp) {
    double qx[G][DIM], qm[G], qs[G];
#pragma unroll
    for (int k = 0; k < G; ++k) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) qx[k][d] = qp[d * G + k];
      qm[k] = qp[DIM * G + k];
      qs[k] = qp[(DIM + 1) * G + k];
    }
#pragma unroll
    for (int k = 0; k < G; ++k) {
      double t = me.a;
#pragma unroll
      for (int d = 0; d < DIM; ++d) t = fma(qx[k][d], me.yl[d], t);
      const double e = expo<MODE>(t - qm[k]);
      if constexpr (BARY) {
        // SotSolver.cc:927-933: (w/S) exp_term x, exp_terms > 0 only
        if (MODE != kF64Plain || e > 0.0) {
          const double wt = qs[k] * e;
          acc += wt;
#pragma unroll
          for (int d = 0; d < DIM; ++d) accx[d] = fma(wt, qx[k][d], accx[d]);
        }
      } else if (MODE == kF64Plain) {
        // the reference adds only exp_terms > 0 (SotSolver.cc:292); in plain
        // mode w/S_q can overflow to inf when S_q is subnormal
        if (e > 0.0) acc = fma(qs[k], e, acc);
      } else {
        acc = fma(qs[k], e, acc);
      }
    }
  };
  // Value bound (VB): the pair (q, j) adds nu_j e^{t_qj - M_q} w_q / S_q
  // to target j's mass; it is skipped when that is below 2^-theta w_q, i.e.
  // when psi~_j - |x_q - y_j|^2/2 < V_q = (M_q + log S_q - theta)/c - |x'_q|^2/2
  // (psi~_j = psi_j + eps ln nu_j), so a target loses less than 2^-theta of
  // mass.  a.vbin / a.vrow hold the smallest V_q of each cell bin / row; the
  // wave's largest psi~ and its box widened by a.rq (the largest distance of a
  // cell's q-points from its center) bound the rest.  (A test per cell cost a
  // quarter of a kept cell's work and was dropped.)
  VBound vb;
  if constexpr (VB) {
    double yy = 0.0;
#pragma unroll
    for (int d = 0; d < DIM; ++d) yy = fma(me.yl[d], me.yl[d], yy);
    const double pt = (me.a + 0.5 * yy * a.inv_coef) * a.inv_coef + a.eps * log(me.nu);
    vb.auni = wave_max(active ? pt : -INFINITY);
    vb.bbin = a.vbin;
    vb.brow = a.vrow;
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      vb.lo[d] = lo[d] - a.rq;
      vb.hi[d] = hi[d] + a.rq;
    }
  }
  auto body = [&](int c0, int c1, auto tag) {
    for (int c = c0; c < c1; ++c) {
      cdouble *cp = ccr + (int64_t)c * 4;
      if constexpr (decltype(tag)::value) {
        double cv[DIM];
#pragma unroll
        for (int d = 0; d < DIM; ++d) cv[d] = cp[d];
        if (!in_set<DIM>(me.y, cv, a.smax)) continue;
      }
      cdouble *qp = qchk + (int64_t)c * nq * (DIM + 2);
      for (int k0 = 0; k0 < nq; k0 += G) chunk(qp + (int64_t)k0 * (DIM + 2));
    }
  };
  auto cells = [&](auto &&f) {
    if constexpr (VB) {
      for_rows_vb<DIM>(a.cg, a.cstart, lo, hi, a.Rq, a.s_in, seg, a.K2, a.dense != 0, vb, f);
    } else if (a.dense) {
      const int64_t c0 = a.n_cells * seg / a.K2, c1 = a.n_cells * (seg + 1) / a.K2;
      f((int)c0, (int)c1, Inside{});
    } else {
      for_rows<DIM>(a.cg, a.cstart, lo, hi, a.Rq, a.s_in, seg, a.K2, f);
    }
  };
  cells(body);
  if (active) {
    a.Gpart[(int64_t)seg * a.N + j] = acc;
    if constexpr (BARY) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) a.Bpart[((int64_t)seg * DIM + d) * a.N + j] = accx[d];
    }
  }
}

// ---------------------------------------------------------------------------
// k_gfin: g_j = nu_j * sum_seg G - nu_j, scattered to caller order
// ---------------------------------------------------------------------------

template <int DIM>
__device__ __forceinline__ void gfin_body(const GFArgs &a, int64_t blk) {
  const int64_t j = blk * 256 + threadIdx.x;
  if (j >= a.N) return;
  const double s = fold_partials(a.Gpart + j, a.N, a.K2, 0.0, FoldAdd{});
  const double nu = a.trec_m[j * kRec + 1 + 2 * DIM];
  a.grad[a.tm2o[j]] = nu * s - nu;
}

// Barycenters (SotSolver.cc:957-963): sum_seg of the first moments plus the
// shift origin times the plan's mass, scattered to caller order.
template <int DIM>
__device__ __forceinline__ void bfin_body(const double *Gpart, const double *Bpart,
                                          const int32_t *tm2o, const double *o, int64_t N,
                                          int K2, double *bary, int64_t blk) {
  const int64_t j = blk * 256 + threadIdx.x;
  if (j >= N) return;
  const double g = fold_partials(Gpart + j, N, K2, 0.0, FoldAdd{});
  double b[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) b[d] = fold_partials(Bpart + d * N + j, DIM * N, K2, 0.0, FoldAdd{});
  const int64_t jo = tm2o[j];
#pragma unroll
  for (int d = 0; d < DIM; ++d) bary[jo * DIM + d] = fma(o[d], g, b[d]);
}

// Fixed-order totals: out[0] = sum F_q - sum psi.nu, out[1] = sum C_q,
// *ucount = number of q-points whose FP32 sum lost its range.

__device__ __forceinline__ void final_body(const FinArgs &a, double *sh) {
  double f = 0.0, c = 0.0, u = 0.0, p = 0.0;
  for (int64_t i = threadIdx.x; i < a.nF; i += 256) {
    f += a.Fpart[i];
    c += a.Cpart[i];
    u += a.Upart[i];
  }
  for (int64_t i = threadIdx.x; i < a.nP; i += 256) p += a.PNpart[i];
  const double fs = block_sum_256(f, sh);
  const double cs = block_sum_256(c, sh);
  const double us = block_sum_256(u, sh);
  const double ps = block_sum_256(p, sh);
  if (threadIdx.x == 0) {
    a.out[0] = fs - ps;
    a.out[1] = cs;
    *a.ucount = us;
  }
}


}  // namespace sdot
