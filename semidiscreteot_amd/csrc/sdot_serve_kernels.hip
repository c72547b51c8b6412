// sdot_serve_kernels.hip -- batched kernels of the serving path: one launch
// advances every active solve.  Each kernel walks a device work list (item ->
// solve by binary search over per-solve item offsets) and runs the same bodies
// as the single-solve kernels (sdot_device.hpp) on the solve's own argument
// blocks, so a solve's arithmetic and reduction order do not depend on which
// other solves share the launch.  The grid is persistent: it is fixed when the
// launch is recorded and loops over however many items the current batch has,
// which lets the host replay one captured graph for batches of any size.
//
// L-BFGS vector work stays on the device (one 1024-thread workgroup per solve,
// fixed-order reductions): the trial step x = x0 + alpha p, the history update
// and two-loop recursion (m = 5, H0 = I; the in-tree text of deal.II's
// SolverBFGS is core/wasserstein_solver_bfgs.h:402-491), the first trial step
// of a line search, and the slopes/stopping statistics the host-side line
// search (sdot_bfgs.hpp LineSearchSM) consumes.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "sdot_device.hpp"
#include "sdot_serve.hpp"

namespace sdot {

__device__ __forceinline__ int find_item(const int64_t *__restrict__ base, int n, int64_t item) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (base[mid] <= item) lo = mid; else hi = mid;
  }
  return lo;
}

// Dynamic item scheduling of the one-wave sweeps.  Workgroups of a
// persistent grid are dispatched round-robin over the 8 XCDs, so workgroup b
// sits on XCD b mod 8; it takes the items i = b (mod 8), in order, from that
// XCD's counter.  Items of equal residue share segments (item counts of every
// solve are multiples of 8), so an XCD keeps streaming its own eighth of the
// records, and a workgroup that finishes early takes more items instead of
// idling at the end of a static share.  Which workgroup runs an item does not
// change what the item computes.
template <class F>
__device__ __forceinline__ void for_items_xcd(int *__restrict__ ctr, int64_t total, F &&f) {
  const int x = blockIdx.x & 7;
  if (x >= total) return;  // (also: an empty list costs no atomics)
  for (;;) {
    int k = 0;
    if (threadIdx.x == 0) k = atomicAdd(ctr + x, 1);
    k = __builtin_amdgcn_readfirstlane(k);
    const int64_t it = x + 8 * (int64_t)k;
    if (it >= total) break;
    f(it);
  }
}

// ---------------------------------------------------------------------------
// evaluation work lists
// ---------------------------------------------------------------------------
template <int DIM>
__global__ __launch_bounds__(256) void kl_prep(const SlotDesc *__restrict__ desc,
                                               const int64_t *__restrict__ base,
                                               const int *__restrict__ nact) {
  __shared__ double sh[256];
  const int n = *nact;
  const int64_t total = base[n];
  for (int64_t it = blockIdx.x; it < total; it += gridDim.x) {
    const int s = find_item(base, n, it);
    const PrepArgs a = desc[s].prep;
    prep_body<DIM>(a, it - base[s], sh);
  }
}

template <int DIM, int G, int MODE, int PASS, bool VB = false>
__global__ __launch_bounds__(64) void kl_pass1(const SlotDesc *__restrict__ desc,
                                               const int64_t *__restrict__ base,
                                               const int *__restrict__ nact,
                                               int *__restrict__ ctr) {
  const int n = *nact;
  for_items_xcd(ctr, base[n], [&](int64_t it) {
    const int s = find_item(base, n, it);
    const P1Args a = desc[s].p1;
    const int64_t loc = it - base[s];
    const int64_t blk = loc / a.K1;
    pass1_body<DIM, G, MODE, PASS, VB>(a, blk, (int)(loc - blk * a.K1));
  });
}

// Value-bound reductions: target bins and rows after kl_prep, cells, cell
// bins and rows after kl_qfin.  An item is 4 bins (one wave each) or 256 rows.
template <int K>
__global__ __launch_bounds__(256) void kl_vbred(const SlotDesc *__restrict__ desc,
                                                const int64_t *__restrict__ base,
                                                const int *__restrict__ nact) {
  const int n = *nact;
  const int64_t total = base[n];
  for (int64_t it = blockIdx.x; it < total; it += gridDim.x) {
    const int s = find_item(base, n, it);
    const VBArgs &v = desc[s].vb;
    const int64_t i = (it - base[s]) * 256 + threadIdx.x;
    const int64_t b = (it - base[s]) * 4 + threadIdx.x / kWave;
    if constexpr (K == kLTbin) bin_reduce_body<true>(v.psit_r, v.tstart, v.ntb, v.pbin, b);
    if constexpr (K == kLTrow) row_reduce_body<true>(v.pbin, v.ntr, v.tnx, v.prow, i);
    if constexpr (K == kLCbin) bin_reduce_body<false>(v.vcell, v.cstart, v.ncb, v.vbin, b);
    if constexpr (K == kLCrow) row_reduce_body<false>(v.vbin, v.ncr, v.cnx, v.vrow, i);
  }
}

template <int DIM>
__global__ __launch_bounds__(256) void kl_cellv(const SlotDesc *__restrict__ desc,
                                                const int64_t *__restrict__ base,
                                                const int *__restrict__ nact) {
  const int n = *nact;
  const int64_t total = base[n];
  for (int64_t it = blockIdx.x; it < total; it += gridDim.x) {
    const int s = find_item(base, n, it);
    const VBArgs &v = desc[s].vb;
    cellv_body<DIM>(v.qrec_r, v.n_cells, v.nq, v.theta, v.inv_coef, v.log2_units != 0, v.vcell,
                    (it - base[s]) * 256 + threadIdx.x);
  }
}

template <int DIM, int MODE>
__global__ __launch_bounds__(256) void kl_qfin(const SlotDesc *__restrict__ desc,
                                               const int64_t *__restrict__ base,
                                               const int *__restrict__ nact) {
  __shared__ double sh[256];
  const int n = *nact;
  const int64_t total = base[n];
  for (int64_t it = blockIdx.x; it < total; it += gridDim.x) {
    const int s = find_item(base, n, it);
    const QFArgs a = desc[s].qf;
    qfin_body<DIM, MODE>(a, it - base[s], sh);
  }
}

template <int DIM, int MODE, int G, bool VB = false>
__global__ __launch_bounds__(64) void kl_grad(const SlotDesc *__restrict__ desc,
                                              const int64_t *__restrict__ base,
                                              const int *__restrict__ nact,
                                              int *__restrict__ ctr) {
  const int n = *nact;
  for_items_xcd(ctr, base[n], [&](int64_t it) {
    const int s = find_item(base, n, it);
    const GArgs a = desc[s].gr;
    const int64_t loc = it - base[s];
    const int64_t blk = loc / a.K2;
    grad_body<DIM, MODE, G, false, VB>(a, blk, (int)(loc - blk * a.K2));
  });
}

template <int DIM>
__global__ __launch_bounds__(256) void kl_gfin(const SlotDesc *__restrict__ desc,
                                               const int64_t *__restrict__ base,
                                               const int *__restrict__ nact) {
  const int n = *nact;
  const int64_t total = base[n];
  for (int64_t it = blockIdx.x; it < total; it += gridDim.x) {
    const int s = find_item(base, n, it);
    const GFArgs a = desc[s].gf;
    gfin_body<DIM>(a, it - base[s]);
  }
}

__global__ __launch_bounds__(256) void kl_final(const SlotDesc *__restrict__ desc,
                                                const int *__restrict__ nact) {
  __shared__ double sh[256];
  const int n = *nact;
  for (int s = blockIdx.x; s < n; s += gridDim.x) {
    const FinArgs a = desc[s].fin;
    final_body(a, sh);
  }
}

// ---------------------------------------------------------------------------
// per-solve L-BFGS vector work (one 1024-thread block per solve)
// ---------------------------------------------------------------------------
constexpr int kBT = 1024;

__device__ __forceinline__ double bsum(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int w = kBT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}
__device__ __forceinline__ double bmax(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
#pragma unroll
  for (int w = kBT / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double bdot(const double *a, const double *b, int64_t n, double *sh) {
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kBT) s += a[i] * b[i];
  return bsum(s, sh);
}

// actions: kActTrial  x = x0 + alpha p
//          kActIter   (history update if has_prev) + two-loop direction + first
//                     trial step of the line search
__global__ __launch_bounds__(kBT) void kl_action(const SlotVecs *__restrict__ vecs,
                                                 const SlotIn *__restrict__ in,
                                                 SlotOut *__restrict__ out,
                                                 const int *__restrict__ nact,
                                                 int *__restrict__ ctr) {
  __shared__ double sh[kBT];
  // the step's item counters start at zero (a kernel, not a memset node, so
  // the step can be replayed from a captured graph)
  if (blockIdx.x == 0 && threadIdx.x < kNumQueues * 8) ctr[threadIdx.x] = 0;
  const int n = *nact;
  for (int s = blockIdx.x; s < n; s += gridDim.x) {
    const SlotVecs v = vecs[s];
    const SlotIn c = in[s];
    const int64_t N = v.N;
    if (c.action == kActTrial) {
      for (int64_t i = threadIdx.x; i < N; i += kBT) v.x[i] = v.x0[i] + c.alpha * v.p[i];
      continue;
    }
    if (c.action != kActIter) continue;
    SlotHist &h = *v.hist;
    int count = h.count, newest = h.newest;
    if (c.has_prev) {
      // s_k = x - x0, y_k = g - g(x0); stored only when s.y > 0
      double cs = 0.0;
      for (int64_t i = threadIdx.x; i < N; i += kBT) cs += (v.x[i] - v.x0[i]) * (v.g[i] - v.gv[i]);
      const double curv = bsum(cs, sh);
      if (curv > 0.0) {
        const int slot = (newest + 1) % kHist;
        for (int64_t i = threadIdx.x; i < N; i += kBT) {
          v.S[slot][i] = v.x[i] - v.x0[i];
          v.Y[slot][i] = v.g[i] - v.gv[i];
        }
        __syncthreads();
        if (threadIdx.x == 0) h.rho[slot] = 1.0 / curv;
        newest = slot;
        count = count < kHist ? count + 1 : kHist;
      }
    }
    // two-loop recursion (newest first), p = -H g
    for (int64_t i = threadIdx.x; i < N; i += kBT) v.p[i] = v.g[i];
    __syncthreads();
    double c1[kHist];
    for (int i = 0; i < count; ++i) {
      const int sl = ((newest - i) % kHist + kHist) % kHist;
      c1[i] = h.rho[sl] * bdot(v.S[sl], v.p, N, sh);
      for (int64_t t = threadIdx.x; t < N; t += kBT) v.p[t] += -c1[i] * v.Y[sl][t];
      __syncthreads();
    }
    if (c.h0 > 0.0) {
      for (int64_t t = threadIdx.x; t < N; t += kBT) v.p[t] *= c.h0 / v.nu[t];
      __syncthreads();
    }
    for (int i = count - 1; i >= 0; --i) {
      const int sl = ((newest - i) % kHist + kHist) % kHist;
      const double c2 = h.rho[sl] * bdot(v.Y[sl], v.p, N, sh);
      const double cc = c1[i] - c2;
      for (int64_t t = threadIdx.x; t < N; t += kBT) v.p[t] += cc * v.S[sl][t];
      __syncthreads();
    }
    for (int64_t i = threadIdx.x; i < N; i += kBT) v.p[i] = -v.p[i];
    __syncthreads();
    const double g0 = bdot(v.g, v.p, N, sh);
    // first trial step (wasserstein_solver_bfgs.h:335-341); 1 in the
    // noise-aware regime (entered before, or |g0| < noise now)
    const double df = fmax(c.f_prev - c.f, 100.0 * 2.220446049250313e-16);
    const double a1 = (c.first || c.surrogate || fabs(g0) < c.noise)
                          ? 1.0 : fmin(1.0, -1.01 * 2.0 * df / g0);
    for (int64_t i = threadIdx.x; i < N; i += kBT) {
      v.x0[i] = v.x[i];
      v.gv[i] = v.g[i];
      v.x[i] = v.x[i] + a1 * v.p[i];
    }
    if (threadIdx.x == 0) {
      out[s].g0 = g0;
      out[s].a1 = a1;
      h.count = count;
      h.newest = newest;
    }
    __syncthreads();
  }
}

// slope along p and stopping statistics of the gradient just evaluated
__global__ __launch_bounds__(kBT) void kl_dots(const SlotVecs *__restrict__ vecs,
                                               const SlotIn *__restrict__ in,
                                               SlotOut *__restrict__ out,
                                               const int *__restrict__ nact) {
  __shared__ double sh[kBT];
  const int n = *nact;
  for (int s = blockIdx.x; s < n; s += gridDim.x) {
    const SlotVecs v = vecs[s];
    const double tol = in[s].tol;
    const int64_t N = v.N;
    double gp = 0.0, l1 = 0.0, cw = -INFINITY, rel = 0.0;
    for (int64_t i = threadIdx.x; i < N; i += kBT) {
      const double g = v.g[i];
      gp += g * v.p[i];
      l1 += fabs(g);
      cw = fmax(cw, fabs(g) - v.nu[i] * tol);
      rel = fmax(rel, fabs(g) / v.nu[i]);
    }
    const double a = bsum(gp, sh);
    const double b = bsum(l1, sh);
    const double c = bmax(cw, sh);
    const double d = bmax(rel, sh);
    if (threadIdx.x == 0) {
      out[s].gp = a;
      out[s].l1 = b;
      out[s].cw = c;
      out[s].relerr = d;
    }
  }
}

// ---------------------------------------------------------------------------
// launcher: one step for the current batch
// ---------------------------------------------------------------------------
template <int DIM, int G, int GS>
static void step_t(const ServeDev &sd, hipStream_t s) {
  const unsigned gp = sd.grid_pass, gg = sd.grid_grad, gb = sd.grid_b, gs = sd.grid_slots;
  const int64_t *b = sd.base;
  const int64_t L = sd.bstride;
  hipLaunchKernelGGL(kl_action, dim3(gs), dim3(kBT), 0, s, sd.vecs, sd.in, sd.out, sd.nact, sd.ctr);
  hipLaunchKernelGGL((kl_prep<DIM>), dim3(gb), dim3(256), 0, s, sd.desc, b + kLPrep * L, sd.nact);
  hipLaunchKernelGGL(kl_vbred<kLTbin>, dim3(gb), dim3(256), 0, s, sd.desc, b + kLTbin * L, sd.nact);
  hipLaunchKernelGGL(kl_vbred<kLTrow>, dim3(gb), dim3(256), 0, s, sd.desc, b + kLTrow * L, sd.nact);
  hipLaunchKernelGGL((kl_pass1<DIM, G, kF32Lse, 2>), dim3(gp), dim3(64), 0, s, sd.desc,
                     b + kLPass32 * L, sd.nact, sd.ctr + kQPass2F32 * 8);
  hipLaunchKernelGGL((kl_pass1<DIM, GS, kF32Lse, 2, true>), dim3(gp), dim3(64), 0, s, sd.desc,
                     b + kLPass32V * L, sd.nact, sd.ctr + kQPass2F32V * 8);
  hipLaunchKernelGGL((kl_pass1<DIM, G, kF64Lse, 0>), dim3(gp), dim3(64), 0, s, sd.desc,
                     b + kLPass64 * L, sd.nact, sd.ctr + kQPass0F64 * 8);
  hipLaunchKernelGGL((kl_pass1<DIM, G, kF64Lse, 1>), dim3(gp), dim3(64), 0, s, sd.desc,
                     b + kLPass64 * L, sd.nact, sd.ctr + kQPass1F64 * 8);
  hipLaunchKernelGGL((kl_qfin<DIM, kF32Lse>), dim3(gb), dim3(256), 0, s, sd.desc,
                     b + kLQfin32 * L, sd.nact);
  hipLaunchKernelGGL((kl_qfin<DIM, kF64Lse>), dim3(gb), dim3(256), 0, s, sd.desc,
                     b + kLQfin64 * L, sd.nact);
  hipLaunchKernelGGL((kl_cellv<DIM>), dim3(gb), dim3(256), 0, s, sd.desc, b + kLCell * L, sd.nact);
  hipLaunchKernelGGL(kl_vbred<kLCbin>, dim3(gb), dim3(256), 0, s, sd.desc, b + kLCbin * L, sd.nact);
  hipLaunchKernelGGL(kl_vbred<kLCrow>, dim3(gb), dim3(256), 0, s, sd.desc, b + kLCrow * L, sd.nact);
  hipLaunchKernelGGL((kl_grad<DIM, kF32Lse, G>), dim3(gg), dim3(64), 0, s, sd.desc,
                     b + kLGrad32 * L, sd.nact, sd.ctr + kQGradF32 * 8);
  hipLaunchKernelGGL((kl_grad<DIM, kF32Lse, GS, true>), dim3(gg), dim3(64), 0, s, sd.desc,
                     b + kLGrad32V * L, sd.nact, sd.ctr + kQGradF32V * 8);
  hipLaunchKernelGGL((kl_grad<DIM, kF64Lse, G>), dim3(gg), dim3(64), 0, s, sd.desc,
                     b + kLGrad64 * L, sd.nact, sd.ctr + kQGradF64 * 8);
  hipLaunchKernelGGL((kl_gfin<DIM>), dim3(gb), dim3(256), 0, s, sd.desc, b + kLGfin * L, sd.nact);
  hipLaunchKernelGGL(kl_final, dim3(gs), dim3(256), 0, s, sd.desc, sd.nact);
  hipLaunchKernelGGL(kl_dots, dim3(gs), dim3(kBT), 0, s, sd.vecs, sd.in, sd.out, sd.nact);
}

// (G, Gs) of a source: G is the largest divisor of nq up to 9 and Gs the
// largest up to 4; these are all the pairs nq = 1 .. 4096 produce.
#define SDOT_SERVE_GROUPINGS(X, D)                                                             \
  X(D, 1, 1) X(D, 2, 2) X(D, 3, 3) X(D, 4, 4) X(D, 5, 1) X(D, 5, 2) X(D, 5, 3) X(D, 5, 4)      \
  X(D, 6, 3) X(D, 6, 4) X(D, 7, 1) X(D, 7, 2) X(D, 7, 3) X(D, 7, 4) X(D, 8, 4) X(D, 9, 3)      \
  X(D, 9, 4)

// Workgroups of `kernel` resident per CU (occupancy), at least 1.
template <class K>
static int resident(K kernel, int threads) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) != hipSuccess) n = 1;
  return n > 0 ? n : 1;
}

template <int DIM, int G, int GS>
static void grids_t(unsigned cus, ServeDev &sd) {
  // the smallest occupancy of the kernels sharing a grid size, so that every
  // persistent workgroup is resident from the start
  int pass = resident(kl_pass1<DIM, G, kF32Lse, 2>, 64);
  pass = std::min(pass, resident(kl_pass1<DIM, GS, kF32Lse, 2, true>, 64));
  pass = std::min(pass, resident(kl_pass1<DIM, G, kF64Lse, 0>, 64));
  pass = std::min(pass, resident(kl_pass1<DIM, G, kF64Lse, 1>, 64));
  int grad = std::min(resident(kl_grad<DIM, kF32Lse, G>, 64), resident(kl_grad<DIM, kF64Lse, G>, 64));
  grad = std::min(grad, resident(kl_grad<DIM, kF32Lse, GS, true>, 64));
  int b = std::min(resident(kl_prep<DIM>, 256), resident(kl_qfin<DIM, kF32Lse>, 256));
  b = std::min(b, std::min(resident(kl_qfin<DIM, kF64Lse>, 256), resident(kl_gfin<DIM>, 256)));
  b = std::min(b, resident(kl_cellv<DIM>, 256));
  // multiples of 8 keep a persistent workgroup's items on one XCD's segments
  sd.grid_pass = cus * (unsigned)pass;
  sd.grid_grad = cus * (unsigned)grad;
  sd.grid_b = cus * (unsigned)b;
}

hipError_t serve_grid_sizes(int dim, int G, int Gs, int device, ServeDev &sd) {
  int cus = 0;
  hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) return e;
  cus = std::max(8, cus - cus % 8);
#define SDOT_GRIDS(D, GG, GGS)                                   \
  if (dim == D && G == GG && Gs == GGS) {                        \
    grids_t<D, GG, GGS>((unsigned)cus, sd);                      \
    return hipSuccess;                                           \
  }
  SDOT_SERVE_GROUPINGS(SDOT_GRIDS, 2)
  SDOT_SERVE_GROUPINGS(SDOT_GRIDS, 3)
#undef SDOT_GRIDS
  return hipErrorInvalidValue;
}

hipError_t launch_serve_step(int dim, int G, int Gs, const ServeDev &sd, hipStream_t s) {
#define SDOT_STEP(D, GG, GGS)                                    \
  if (dim == D && G == GG && Gs == GGS) {                        \
    step_t<D, GG, GGS>(sd, s);                                   \
    return hipGetLastError();                                    \
  }
  SDOT_SERVE_GROUPINGS(SDOT_STEP, 2)
  SDOT_SERVE_GROUPINGS(SDOT_STEP, 3)
#undef SDOT_STEP
  return hipErrorInvalidValue;
}

}  // namespace sdot
