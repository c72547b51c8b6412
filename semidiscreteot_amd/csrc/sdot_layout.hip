// sdot_layout.hip -- the target layout built on the device.
//
// The q-major sweeps stream the targets in row-major bin order and the
// gradient pass gives each wave 64 Morton-consecutive targets; both orders,
// the bin offsets and the waves' bounding boxes are built here from the points
// in caller order, in HBM:
//   k_lay_keys    per target: row-major bin index, Morton key, its own index
//   radix sorts   stable (LSD) sorts of (bin, index) and (Morton key, index)
//   k_lay_scatter inverse permutations; points in Morton order (padded to 4)
//   k_lay_start   bin offsets: first sorted position of every bin
//   k_lay_box     per wave of 64 Morton-consecutive targets: min / max corner
// Stable sorts keep the targets of one bin in caller order, so the record
// streams (and every sum over them) are the same as with any stable sort.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include <hipcub/device/device_radix_sort.hpp>

#include "sdot_internal.hpp"

namespace sdot {

namespace {

template <int DIM>
__global__ __launch_bounds__(256) void k_lay_keys(const double *__restrict__ y, int64_t n, Grid g,
                                                  double lo0, double lo1, double lo2, double hi0,
                                                  double hi1, double hi2, uint32_t *bin,
                                                  uint64_t *mkey, int32_t *idx) {
  const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (j >= n) return;
  double p[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) p[d] = y[j * DIM + d];
  const double lo[3] = {lo0, lo1, lo2}, hi[3] = {hi0, hi1, hi2};
  bin[j] = (uint32_t)bin_linear(g, p, DIM);
  mkey[j] = morton_key(p, lo, hi, DIM);
  idx[j] = (int32_t)j;
}

template <int DIM>
__global__ __launch_bounds__(256) void k_lay_scatter(const double *__restrict__ y, int64_t n,
                                                     const int32_t *__restrict__ r2o,
                                                     const int32_t *__restrict__ m2o,
                                                     int32_t *pos_r, int32_t *pos_m,
                                                     double *ty_m) {
  const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (k >= n) return;
  pos_r[r2o[k]] = (int32_t)k;
  const int32_t j = m2o[k];
  pos_m[j] = (int32_t)k;
  if (ty_m) {
#pragma unroll
    for (int d = 0; d < 4; ++d) ty_m[k * 4 + d] = d < DIM ? y[(int64_t)j * DIM + d] : 0.0;
  }
}

// tstart[b] = number of targets in bins < b (b = 0 .. n_bins).
__global__ __launch_bounds__(256) void k_lay_start(const uint32_t *__restrict__ sorted_bin,
                                                   int64_t n, int64_t n_bins, int32_t *tstart) {
  const int64_t b = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (b > n_bins) return;
  int64_t lo = 0, hi = n;  // first position whose bin is >= b
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)sorted_bin[mid] < b) lo = mid + 1; else hi = mid;
  }
  tstart[b] = (int32_t)lo;
}

// One wave per box of 64 Morton-consecutive targets.
template <int DIM>
__global__ __launch_bounds__(256) void k_lay_box(const double *__restrict__ y, int64_t n,
                                                 const int32_t *__restrict__ m2o,
                                                 int64_t n_waves, double *twbox) {
  const int64_t w = (int64_t)blockIdx.x * 4 + threadIdx.x / kWave;
  if (w >= n_waves) return;  // whole waves leave together
  const int64_t k = w * kWave + (threadIdx.x & (kWave - 1));
  double mn[DIM], mx[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const double v = k < n ? y[(int64_t)m2o[k] * DIM + d] : 0.0;
    mn[d] = k < n ? v : INFINITY;
    mx[d] = k < n ? v : -INFINITY;
  }
#pragma unroll
  for (int m = kWave / 2; m >= 1; m >>= 1)
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      mn[d] = fmin(mn[d], __shfl_xor(mn[d], m));
      mx[d] = fmax(mx[d], __shfl_xor(mx[d], m));
    }
  if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      twbox[w * 2 * DIM + d] = mn[d];
      twbox[w * 2 * DIM + DIM + d] = mx[d];
    }
  }
}

// Bits that hold every value below `count`.
int key_bits(uint64_t count) {
  int b = 1;
  while (b < 64 && (count >> b) != 0) ++b;
  return b;
}

// Scratch layout: bin keys in/out, Morton keys in/out, indices, a row-order
// index buffer (when the caller keeps no r2o), then the sorts' temporary storage.
struct LayScratch {
  size_t bin_in, bin_out, mk_in, mk_out, idx, r2o, tmp, tmp_bytes, total;
};

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

LayScratch lay_scratch(const TargetShape &S) {
  const size_t n = (size_t)S.n;
  size_t t_bin = 0, t_mk = 0;
  const hipError_t e1 = hipcub::DeviceRadixSort::SortPairs(
      nullptr, t_bin, (const uint32_t *)nullptr, (uint32_t *)nullptr, (const int32_t *)nullptr,
      (int32_t *)nullptr, (int)n, 0, key_bits((uint64_t)S.n_bins));
  const hipError_t e2 = hipcub::DeviceRadixSort::SortPairs(
      nullptr, t_mk, (const uint64_t *)nullptr, (uint64_t *)nullptr, (const int32_t *)nullptr,
      (int32_t *)nullptr, (int)n, 0, 64);
  LayScratch L{};
  if (e1 != hipSuccess || e2 != hipSuccess) {  // no size: no caller's scratch fits
    L.total = SIZE_MAX;
    return L;
  }
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += align256(bytes);
    return o;
  };
  L.bin_in = take(n * 4);
  L.bin_out = take(n * 4);
  L.mk_in = take(n * 8);
  L.mk_out = take(n * 8);
  L.idx = take(n * 4);
  L.r2o = take(n * 4);
  L.tmp_bytes = std::max(t_bin, t_mk);
  L.tmp = take(L.tmp_bytes);
  L.total = off;
  return L;
}

}  // namespace

size_t layout_scratch_bytes(const TargetShape &S) { return lay_scratch(S).total; }

hipError_t device_target_layout(int dim, const TargetShape &S, const double *d_y, int32_t *pos_r,
                                int32_t *r2o, int32_t *pos_m, int32_t *m2o, int32_t *tstart,
                                double *twbox, double *ty_m, void *scratch, size_t scratch_bytes,
                                hipStream_t s) {
  if (dim != 2 && dim != 3) return hipErrorInvalidValue;
  if (S.n < 1 || S.n >= (int64_t)1 << 31 || S.n_bins < 1 || S.n_bins >= (int64_t)1 << 31)
    return hipErrorInvalidValue;
  const LayScratch L = lay_scratch(S);
  if (!scratch || scratch_bytes < L.total || !pos_r || !pos_m || !m2o || !tstart || !twbox)
    return hipErrorInvalidValue;
  char *base = static_cast<char *>(scratch);
  uint32_t *bin_in = (uint32_t *)(base + L.bin_in), *bin_out = (uint32_t *)(base + L.bin_out);
  uint64_t *mk_in = (uint64_t *)(base + L.mk_in), *mk_out = (uint64_t *)(base + L.mk_out);
  int32_t *idx = (int32_t *)(base + L.idx);
  int32_t *rows = r2o ? r2o : (int32_t *)(base + L.r2o);
  void *tmp = base + L.tmp;
  const int64_t n = S.n;
  const unsigned nb = (unsigned)((n + 255) / 256);
  if (dim == 2)
    hipLaunchKernelGGL(k_lay_keys<2>, dim3(nb), dim3(256), 0, s, d_y, n, S.tg, S.lo[0], S.lo[1],
                       S.lo[2], S.hi[0], S.hi[1], S.hi[2], bin_in, mk_in, idx);
  else
    hipLaunchKernelGGL(k_lay_keys<3>, dim3(nb), dim3(256), 0, s, d_y, n, S.tg, S.lo[0], S.lo[1],
                       S.lo[2], S.hi[0], S.hi[1], S.hi[2], bin_in, mk_in, idx);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  size_t tb = L.tmp_bytes;
  e = hipcub::DeviceRadixSort::SortPairs(tmp, tb, bin_in, bin_out, idx, rows, (int)n, 0,
                                         key_bits((uint64_t)S.n_bins), s);
  if (e != hipSuccess) return e;
  tb = L.tmp_bytes;
  e = hipcub::DeviceRadixSort::SortPairs(tmp, tb, mk_in, mk_out, idx, m2o, (int)n, 0, 64, s);
  if (e != hipSuccess) return e;
  if (dim == 2)
    hipLaunchKernelGGL(k_lay_scatter<2>, dim3(nb), dim3(256), 0, s, d_y, n, rows, m2o, pos_r,
                       pos_m, ty_m);
  else
    hipLaunchKernelGGL(k_lay_scatter<3>, dim3(nb), dim3(256), 0, s, d_y, n, rows, m2o, pos_r,
                       pos_m, ty_m);
  hipLaunchKernelGGL(k_lay_start, dim3((unsigned)((S.n_bins + 1 + 255) / 256)), dim3(256), 0, s,
                     bin_out, n, S.n_bins, tstart);
  const unsigned nbw = (unsigned)((S.n_twaves + 3) / 4);
  if (dim == 2)
    hipLaunchKernelGGL(k_lay_box<2>, dim3(nbw), dim3(256), 0, s, d_y, n, m2o, S.n_twaves, twbox);
  else
    hipLaunchKernelGGL(k_lay_box<3>, dim3(nbw), dim3(256), 0, s, d_y, n, m2o, S.n_twaves, twbox);
  return hipGetLastError();
}

}  // namespace sdot
