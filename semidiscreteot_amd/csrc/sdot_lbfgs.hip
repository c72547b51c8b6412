// sdot_lbfgs.hip -- the two-loop recursion of the L-BFGS driver on device
// vectors, for solves whose potentials are long enough that the recursion
// over the stored (s, y) pairs costs more on one host thread than the
// transfers of g and p (10 dot products and 10 AXPYs per iteration).
//
// Every dot product is a fixed grid of kLbParts partial sums, reduced in the
// same fixed order by every block of the update that consumes it, so the
// direction does not depend on the launch shape of the machine.  The
// coefficient of each dot feeds the next update in device memory: the
// recursion runs without host round trips.
#include <hip/hip_runtime.h>

#include "sdot_internal.hpp"

namespace sdot {
namespace {

constexpr int kLbT = 256;
constexpr int kLbMaxBlocks = 2048;
enum { kLbNone = 0, kLbFirst = 1, kLbSecond = 2 };

__device__ __forceinline__ double lb_block_sum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int w = kLbT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  return sh[0];
}

// part[b] = block b's share of sum_i a[i] b[i] (grid of exactly kLbParts blocks)
__global__ __launch_bounds__(kLbT) void k_lb_dot(const double *__restrict__ a,
                                                 const double *__restrict__ b, int64_t n,
                                                 double *__restrict__ part) {
  __shared__ double sh[kLbT];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * kLbT + threadIdx.x; i < n; i += (int64_t)kLbParts * kLbT)
    s += a[i] * b[i];
  const double r = lb_block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// p <- sign * h0 .* (p + c v), where c comes from the partial sums of the
// dot product just taken:
//   kLbFirst   coef = rho dot (kept for the second loop), c = -coef
//   kLbSecond  c = coef - rho dot
//   kLbNone    c = 0 (v and part unused)
// h0 == nullptr: no scaling; sign is +1 or -1 (exact).
__global__ __launch_bounds__(kLbT) void k_lb_update(double *__restrict__ p,
                                                    const double *__restrict__ v, int64_t n,
                                                    const double *__restrict__ part, double rho,
                                                    double *__restrict__ coef, int mode,
                                                    const double *__restrict__ h0, double sign) {
  __shared__ double sh[kLbT];
  double c = 0.0;
  if (mode != kLbNone) {
    double s = 0.0;
    for (int k = threadIdx.x; k < kLbParts; k += kLbT) s += part[k];
    const double dot = lb_block_sum(s, sh);
    if (mode == kLbFirst) {
      const double cf = rho * dot;
      c = -cf;
      if (blockIdx.x == 0 && threadIdx.x == 0) *coef = cf;
    } else {
      c = *coef - rho * dot;
    }
  }
  const int64_t stride = (int64_t)gridDim.x * kLbT;
  for (int64_t i = (int64_t)blockIdx.x * kLbT + threadIdx.x; i < n; i += stride) {
    double x = p[i];
    if (mode != kLbNone) x += c * v[i];
    if (h0) x *= h0[i];
    p[i] = sign * x;
  }
}

}  // namespace

hipError_t launch_lb_direction(const LbVecs &L, const int *slots, const double *rho, int count,
                               hipStream_t s) {
  if (L.n <= 0) return hipSuccess;
  const int64_t want = (L.n + kLbT - 1) / kLbT;
  const unsigned gu = (unsigned)(want < kLbMaxBlocks ? want : kLbMaxBlocks);
  const size_t n = (size_t)L.n;
  if (count == 0) {
    hipLaunchKernelGGL(k_lb_update, dim3(gu), dim3(kLbT), 0, s, L.p, nullptr, L.n, nullptr, 0.0,
                       nullptr, (int)kLbNone, L.h0, -1.0);
    return hipGetLastError();
  }
  // first loop, newest pair first: coef_i = rho_i s_i.p, p -= coef_i y_i; then p *= h0
  for (int i = 0; i < count; ++i) {
    const double *Si = L.S + (size_t)slots[i] * n, *Yi = L.Y + (size_t)slots[i] * n;
    hipLaunchKernelGGL(k_lb_dot, dim3(kLbParts), dim3(kLbT), 0, s, Si, L.p, L.n, L.part);
    hipLaunchKernelGGL(k_lb_update, dim3(gu), dim3(kLbT), 0, s, L.p, Yi, L.n, L.part, rho[i],
                       L.coef + i, (int)kLbFirst, i == count - 1 ? L.h0 : nullptr, 1.0);
  }
  // second loop, oldest first: p += (coef_i - rho_i y_i.p) s_i; then p = -p
  for (int i = count - 1; i >= 0; --i) {
    const double *Si = L.S + (size_t)slots[i] * n, *Yi = L.Y + (size_t)slots[i] * n;
    hipLaunchKernelGGL(k_lb_dot, dim3(kLbParts), dim3(kLbT), 0, s, Yi, L.p, L.n, L.part);
    hipLaunchKernelGGL(k_lb_update, dim3(gu), dim3(kLbT), 0, s, L.p, Si, L.n, L.part, rho[i],
                       L.coef + i, (int)kLbSecond, nullptr, i == 0 ? -1.0 : 1.0);
  }
  return hipGetLastError();
}

}  // namespace sdot
