// sdot_lbfgs.hip -- the vector arithmetic of the L-BFGS driver on device
// vectors, for solves whose potentials are long enough that its loops over
// host vectors and the transfers of x and g every evaluation cost more than
// the launches (the two-loop recursion alone is 10 dot products and 10
// AXPYs per iteration).
//
// Every reduction is a fixed grid of kLbParts partial values, reduced in one
// fixed order by its consumer (the blocks of the next update in the two-loop
// recursion, or the host), so results do not depend on the machine's launch
// shape.  Inside the recursion each dot product's coefficient feeds the next
// update in device memory: no host round trips.
#include <hip/hip_runtime.h>

#include <cmath>

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
  const double r = sh[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double lb_block_max(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int w = kLbT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
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

// the statistics of launch_lb_stats (grid of exactly kLbParts blocks); the
// minimum is kept as the maximum of -a
__global__ __launch_bounds__(kLbT) void k_lb_stats(const double *__restrict__ a,
                                                   const double *__restrict__ nu, double t,
                                                   int64_t n, double *__restrict__ part) {
  __shared__ double sh[kLbT];
  double l1 = 0.0, l2 = 0.0, cw = -INFINITY, mx = -INFINITY, mn = -INFINITY;
  for (int64_t i = (int64_t)blockIdx.x * kLbT + threadIdx.x; i < n;
       i += (int64_t)kLbParts * kLbT) {
    const double v = a[i], av = fabs(v);
    l1 += av;
    l2 += v * v;
    if (nu) cw = fmax(cw, av - t * nu[i]);
    mx = fmax(mx, v);
    mn = fmax(mn, -v);
  }
  l1 = lb_block_sum(l1, sh);
  l2 = lb_block_sum(l2, sh);
  cw = lb_block_max(cw, sh);
  mx = lb_block_max(mx, sh);
  mn = lb_block_max(mn, sh);
  if (threadIdx.x == 0) {
    part[kLbL1 * kLbParts + blockIdx.x] = l1;
    part[kLbL2 * kLbParts + blockIdx.x] = l2;
    part[kLbCw * kLbParts + blockIdx.x] = cw;
    part[kLbMax * kLbParts + blockIdx.x] = mx;
    part[kLbMin * kLbParts + blockIdx.x] = -mn;
  }
}

__global__ __launch_bounds__(kLbT) void k_lb_lin(double *__restrict__ x,
                                                 const double *__restrict__ x0, double a,
                                                 const double *__restrict__ p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kLbT;
  for (int64_t i = (int64_t)blockIdx.x * kLbT + threadIdx.x; i < n; i += stride)
    x[i] = x0[i] + a * p[i];
}

__global__ __launch_bounds__(kLbT) void k_lb_diff(double *__restrict__ s,
                                                  const double *__restrict__ x,
                                                  double *__restrict__ y,
                                                  const double *__restrict__ g, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kLbT;
  for (int64_t i = (int64_t)blockIdx.x * kLbT + threadIdx.x; i < n; i += stride) {
    s[i] = x[i] - s[i];
    y[i] = g[i] - y[i];
  }
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

unsigned elementwise_grid(int64_t n) {
  const int64_t want = (n + kLbT - 1) / kLbT;
  return (unsigned)(want < 1 ? 1 : (want < kLbMaxBlocks ? want : kLbMaxBlocks));
}

}  // namespace

hipError_t launch_lb_dot(const double *a, const double *b, int64_t n, double *part, hipStream_t s) {
  hipLaunchKernelGGL(k_lb_dot, dim3(kLbParts), dim3(kLbT), 0, s, a, b, n, part);
  return hipGetLastError();
}

hipError_t launch_lb_lin(double *x, const double *x0, double a, const double *p, int64_t n,
                         hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lb_lin, dim3(elementwise_grid(n)), dim3(kLbT), 0, s, x, x0, a, p, n);
  return hipGetLastError();
}

hipError_t launch_lb_diff(double *sv, const double *x, double *y, const double *g, int64_t n,
                          hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_lb_diff, dim3(elementwise_grid(n)), dim3(kLbT), 0, s, sv, x, y, g, n);
  return hipGetLastError();
}

hipError_t launch_lb_stats(const double *a, const double *nu, double t, int64_t n, double *part,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_lb_stats, dim3(kLbParts), dim3(kLbT), 0, s, a, nu, t, n, part);
  return hipGetLastError();
}

hipError_t launch_lb_direction(const LbVecs &L, const int *slots, const double *rho, int count,
                               hipStream_t s) {
  if (L.n <= 0) return hipSuccess;
  if (count < 0 || count > kLbMaxHist) return hipErrorInvalidValue;
  const unsigned gu = elementwise_grid(L.n);
  if (count == 0) {
    hipLaunchKernelGGL(k_lb_update, dim3(gu), dim3(kLbT), 0, s, L.p, nullptr, L.n, nullptr, 0.0,
                       nullptr, (int)kLbNone, L.h0, -1.0);
    return hipGetLastError();
  }
  // first loop, newest pair first: coef_i = rho_i s_i.p, p -= coef_i y_i; then p *= h0
  for (int i = 0; i < count; ++i) {
    const double *Si = L.S[slots[i]], *Yi = L.Y[slots[i]];
    hipLaunchKernelGGL(k_lb_dot, dim3(kLbParts), dim3(kLbT), 0, s, Si, L.p, L.n, L.part);
    hipLaunchKernelGGL(k_lb_update, dim3(gu), dim3(kLbT), 0, s, L.p, Yi, L.n, L.part, rho[i],
                       L.coef + i, (int)kLbFirst, i == count - 1 ? L.h0 : nullptr, 1.0);
  }
  // second loop, oldest first: p += (coef_i - rho_i y_i.p) s_i; then p = -p
  for (int i = count - 1; i >= 0; --i) {
    const double *Si = L.S[slots[i]], *Yi = L.Y[slots[i]];
    hipLaunchKernelGGL(k_lb_dot, dim3(kLbParts), dim3(kLbT), 0, s, Yi, L.p, L.n, L.part);
    hipLaunchKernelGGL(k_lb_update, dim3(gu), dim3(kLbT), 0, s, L.p, Si, L.n, L.part, rho[i],
                       L.coef + i, (int)kLbSecond, nullptr, i == 0 ? -1.0 : 1.0);
  }
  return hipGetLastError();
}

}  // namespace sdot
