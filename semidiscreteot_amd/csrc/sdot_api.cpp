// sdot_api.cpp -- the C ABI (include/sdot.h): solver context, device layout
// of the measures, distance thresholds, the BFGS driver and the stopping test.
// Everything numerical runs in the HIP kernels of sdot_kernels.hip; this file
// only prepares layouts, computes the scalar truncation radius on the host in
// FP64 (so that candidate sets are bit-identical to the FP64 CPU evaluation)
// and drives L-BFGS.
#include <algorithm>
#include <array>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "sdot_bfgs.hpp"
#include "sdot_internal.hpp"

using namespace sdot;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return fail(SDOT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));    \
  } while (0)

template <class T>
hipError_t dalloc(T *&p, size_t n) {
  if (p) {
    (void)hipFree(p);
    p = nullptr;
  }
  if (n == 0) n = 1;
  return hipMalloc((void **)&p, n * sizeof(T));
}

template <class T>
hipError_t upload(T *&p, const std::vector<T> &h) {
  hipError_t e = dalloc(p, h.size());
  if (e != hipSuccess) return e;
  if (h.empty()) return hipSuccess;
  return hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
}

bool all_finite(const double *a, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (!std::isfinite(a[i])) return false;
  return true;
}

// Morton (Z-order) key of a point inside [lo, hi].
// Uniform grid over [lo, hi] with about `per_bin` points per bin.
Grid make_grid(const double *lo, const double *hi, int dim, int64_t n, double per_bin) {
  Grid g{};
  double ext[3] = {1, 1, 1}, maxext = 0;
  for (int d = 0; d < dim; ++d) maxext = std::max(maxext, hi[d] - lo[d]);
  if (!(maxext > 0)) maxext = 1.0;
  double vol = 1.0;
  for (int d = 0; d < dim; ++d) {
    ext[d] = std::max(hi[d] - lo[d], maxext * 1e-6);
    vol *= ext[d];
  }
  double h = std::pow(vol * per_bin / (double)std::max<int64_t>(n, 1), 1.0 / dim);
  auto nbins = [&](double hh) {
    double t = 1;
    for (int d = 0; d < dim; ++d) t *= std::max(1.0, std::ceil(ext[d] / hh));
    return t;
  };
  while (nbins(h) > (double)(1 << 22)) h *= 1.25;
  g.h = h;
  g.inv_h = 1.0 / h;
  for (int d = 0; d < 3; ++d) {
    g.o[d] = d < dim ? lo[d] : 0.0;
    g.n[d] = d < dim ? (int)std::max(1.0, std::ceil(ext[d] / h)) : 1;
  }
  return g;
}

constexpr int kMaxSeg = 128;
constexpr int kProfSets = 128;

}  // namespace

// Largest double s with sqrt(s) <= R (IEEE sqrt is correctly rounded and
// monotone), so that `s <= smax` is the reference's `sqrt(s) <= R`.
double sdot::smax_for(double R) {
  if (std::isnan(R) || R < 0) return -1.0;
  if (std::isinf(R)) return INFINITY;
  double s = R * R;
  if (std::isinf(s)) s = DBL_MAX;
  while (s > 0 && std::sqrt(s) > R) s = std::nextafter(s, 0.0);
  for (;;) {
    const double n = std::nextafter(s, INFINITY);
    if (std::isinf(n) || std::sqrt(n) > R) break;
    s = n;
  }
  return s;
}

int sdot::set_error(int code, const std::string &msg) { return fail(code, msg); }

// Squared radius below which a pair surely passes s <= smax: a relative margin
// of 1e-12, far above the rounding of the FP64 squared distance (a few ulps).
double sdot::inner_s(double smax) {
  if (!(smax > 0.0) || !std::isfinite(smax)) return 0.0;
  return smax * (1.0 - 1e-12);
}

// Largest double s with pow(sqrt(s), 2) <= R*R, the per-q test of the softmax
// refinement (pow(d, 2) is d*d, one rounding).  s -> fl(fl(sqrt(s))^2) is
// monotone, so the test is s <= s2max.
double sdot::s2max_for(double R) {
  if (std::isnan(R)) return -1.0;
  const double T = R * R;
  if (std::isinf(T)) return INFINITY;
  auto ok = [T](double s) {
    const double d = std::sqrt(s);
    return d * d <= T;
  };
  double s = T;
  while (s > 0 && !ok(s)) s = std::nextafter(s, 0.0);
  for (;;) {
    const double n = std::nextafter(s, INFINITY);
    if (std::isinf(n) || !ok(n)) break;
    s = n;
  }
  return s;
}

// Bounding box, coordinate scale and bin grid of a target set: one pass over
// the points.
void sdot::target_shape(int dim, int64_t n, const double *y, TargetShape &S) {
  S.n = n;
  S.n_twaves = (n + kWave - 1) / kWave;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) {
    lo[d] = INFINITY;
    hi[d] = -INFINITY;
  }
  for (int64_t j = 0; j < n; ++j)
    for (int d = 0; d < dim; ++d) {
      lo[d] = std::min(lo[d], y[j * dim + d]);
      hi[d] = std::max(hi[d], y[j * dim + d]);
    }
  S.scale = 0.0;
  for (int d = 0; d < 3; ++d) {
    S.lo[d] = d < dim ? lo[d] : 0.0;
    S.hi[d] = d < dim ? hi[d] : 0.0;
    if (d < dim) S.scale = std::max({S.scale, std::fabs(lo[d]), std::fabs(hi[d])});
  }
  S.tg = make_grid(lo, hi, dim, n, 8.0);
  S.n_bins = (int64_t)S.tg.n[0] * S.tg.n[1] * S.tg.n[2];
}

// The layout of device_target_layout on the host: stable sorts by Morton key
// and by row-major bin.
void sdot::build_target_layout(int dim, int64_t n, const double *y, TargetLayout &L) {
  target_shape(dim, n, y, L.shape);
  const TargetShape &S = L.shape;
  std::vector<uint64_t> key(n);
  for (int64_t j = 0; j < n; ++j) key[j] = morton_key(y + j * dim, S.lo, S.hi, dim);
  L.m2o.resize(n);
  L.pos_m.resize(n);
  std::iota(L.m2o.begin(), L.m2o.end(), 0);
  std::stable_sort(L.m2o.begin(), L.m2o.end(), [&](int32_t a, int32_t b) { return key[a] < key[b]; });
  for (int64_t m = 0; m < n; ++m) L.pos_m[L.m2o[m]] = (int32_t)m;
  std::vector<int> tb(n);
  for (int64_t j = 0; j < n; ++j) tb[j] = bin_linear(S.tg, y + j * dim, dim);
  L.r2o.resize(n);
  L.pos_r.resize(n);
  std::iota(L.r2o.begin(), L.r2o.end(), 0);
  std::stable_sort(L.r2o.begin(), L.r2o.end(), [&](int32_t a, int32_t b) { return tb[a] < tb[b]; });
  for (int64_t r = 0; r < n; ++r) L.pos_r[L.r2o[r]] = (int32_t)r;
  L.tstart.assign(S.n_bins + 1, 0);
  for (int64_t j = 0; j < n; ++j) L.tstart[tb[j] + 1]++;
  for (int64_t b = 0; b < S.n_bins; ++b) L.tstart[b + 1] += L.tstart[b];
}

struct sdot_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  sdot_params params{};
  // members SotSolver keeps between evaluations (SotSolver.h:505-514)
  double C_global = 0.0, global_functional = 0.0, current_threshold = 0.1;
  double covering_radius = 0.0, min_target_density = 0.0;
  // the raw covering radius of the current source and targets (< 0: not known)
  double cover_r0 = -1.0;
  int last_iterations = 0, converged = 0;
  int64_t last_pairs = -1;
  bool have_source = false, have_target = false;
  int force_mode = -1;  // overrides the arithmetic mode (the Hessian runs FP64 LSE)
  bool hist_valid = false;  // device holds psi and M_q of a previous FP32 evaluation
  double hist_R = 0.0;      // its truncation radius
  int hist_G = 0, hist_small = -1, hist_K1 = 0, hist_dense = -1;  // and its work items
  int n_servers = 0;        // serving queues reading this context's source arrays
  double seed_bias = 0.0;   // test hooks (sdot_ctx_create)
  int force_K1 = 0, force_K2 = 0;
  double *noalloc = nullptr;  // this device's no-allocation scratch (kNoAllocBytes)
  int64_t n_tbins = 0;        // bins of the target grid
  Plan plan;
  Dev dev;
  std::vector<double> h_nu;
  std::vector<int64_t> h_cell_r2o;  // row-order cell -> caller cell
  double src_lo[3] = {0, 0, 0}, src_hi[3] = {0, 0, 0};
  double tgt_lo[3] = {0, 0, 0}, tgt_hi[3] = {0, 0, 0};
  double coord_scale = 1.0;
  size_t cap_trec = 0, cap_qrec = 0, cap_M = 0, cap_G = 0, cap_F = 0, cap_P = 0, cap_N = 0;
  size_t cap_vb_t = 0, cap_vb_c = 0;  // value-bound arrays (target side, cell side)
  // profiling: a pool of event sets resolved at synchronisation points
  bool prof_on = false;
  std::vector<std::array<hipEvent_t, 6>> ev_pool;
  int ev_pending = 0;
  sdot_profile prof_acc{};
  std::vector<double> h_tmp, h_tmp2;
};

namespace {

int ctx_check(sdot_ctx *c) {
  if (!c) return fail(SDOT_ERR_INVALID, "null context");
  return SDOT_OK;
}

int resolve_profile(sdot_ctx *c) {
  if (c->ev_pending == 0) return SDOT_OK;
  HIP_TRY(hipStreamSynchronize(c->stream));
  for (int i = 0; i < c->ev_pending; ++i) {
    auto &ev = c->ev_pool[i];
    float t[5] = {0, 0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) HIP_TRY(hipEventElapsedTime(&t[k], ev[k], ev[k + 1]));
    HIP_TRY(hipEventElapsedTime(&t[4], ev[0], ev[5]));
    c->prof_acc.ms_lse_max += t[0];
    c->prof_acc.ms_lse_sum += t[1];
    c->prof_acc.ms_qfin += t[2];
    c->prof_acc.ms_grad += t[3];
    c->prof_acc.ms_total += t[4];
    c->prof_acc.evaluations += 1;
  }
  c->ev_pending = 0;
  return SDOT_OK;
}

// The evaluation plan at radius R: arithmetic mode, candidate rule, grouping
// and segment counts.
int plan_eval(sdot_ctx *c, double R) {
  if (!c->have_source || !c->have_target)
    return fail(SDOT_ERR_STATE, "Invalid measures configuration: source and target must be set");
  Plan &p = c->plan;
  p.eps = c->params.epsilon;
  if (!(p.eps > 0) || !std::isfinite(p.eps)) return fail(SDOT_ERR_INVALID, "epsilon must be > 0");
  const bool lse = c->params.use_log_sum_exp_trick != 0;
  p.mode = !lse ? kF64Plain : (c->params.precision == SDOT_PREC_FP32_LSE ? kF32Lse : kF64Lse);
  if (c->force_mode >= 0) p.mode = c->force_mode;
  // dense when every center-target distance is surely <= R
  double diag2 = 0.0;
  for (int d = 0; d < p.dim; ++d) {
    const double lo = std::min(c->src_lo[d], c->tgt_lo[d]);
    const double hi = std::max(c->src_hi[d], c->tgt_hi[d]);
    diag2 += (hi - lo) * (hi - lo);
  }
  const double diag = std::sqrt(diag2);
  p.smax = smax_for(R);
  p.s_in = inner_s(p.smax);
  p.dense = (R >= diag * (1.0 + 1e-6) + 1e-300) ? 1 : 0;
  p.Rq = R * (1.0 + 1e-9) + 1e-9 * (p.tg.h + p.cg.h + c->coord_scale);
  p.seed_bias = c->seed_bias;
  if (!std::isfinite(p.Rq)) p.dense = std::isnan(R) ? 0 : 1;
  if (std::isnan(R)) p.Rq = 0.0;
  // truncated sweeps run the small grouping (same R, same grouping: the FP32
  // history stays valid)
  // value bound: off in plain mode and with a zero target weight; below 0 the
  // precision's default, which bounds FP32_LSE evaluations where it pays: the
  // radius sqrt(2 bits ln2 eps) beyond which a target's terms fall below
  // 2^-bits (at equal potentials) is at most a fifth of the domain's diagonal,
  // and there are at least 2^32 pairs.  Measured on the 32^3 source (one
  // evaluation, FP32): 50,000 targets, eps 1e-3: 6.7 -> 5.0 ms truncated,
  // 7.5 -> 5.7 ms dense; 20,000 targets, eps 3e-3 (radius 0.26 of the
  // diagonal): 3.2 -> 3.4 ms; 4,096 targets, eps 1e-3: 0.87 -> 0.94 ms.
  if (c->params.prune_bits < 0) {
    double sdiag2 = 0.0;
    for (int d = 0; d < p.dim; ++d)
      sdiag2 += (c->src_hi[d] - c->src_lo[d]) * (c->src_hi[d] - c->src_lo[d]);
    p.prune_bits = (p.mode == kF32Lse &&
                    value_bound_default(p.eps, std::sqrt(sdiag2), R, p.N, p.nQ))
                       ? kPruneBitsF32 : 0.0;
  } else {
    p.prune_bits = c->params.prune_bits;
  }
  if (p.mode == kF64Plain || !(p.nu_min > 0.0) || c->force_mode >= 0) p.prune_bits = 0.0;
  {
    // bounded dense sweeps take the small grouping too: its blocks span half
    // the cells, which tightens the wave's value bound
    const bool small = !p.dense || p.prune_bits > 0;
    const Grouping &gr = (small && p.grp_small.G != p.grp_big.G) ? p.grp_small : p.grp_big;
    p.small = &gr == &p.grp_small ? 1 : 0;
    p.G = gr.G;
    p.gpc = gr.gpc;
    p.n_groups = gr.n_groups;
    p.n_blocks = gr.n_blocks;
  }
  // enough one-wave workgroups for several rounds on the 1,024 SIMDs: the
  // work per wave varies several-fold with the wave's position (candidate
  // balls clipped by the domain), and only later rounds balance it (measured
  // on the 32^3 / 50,000-target workload: 8 -> 32 q-major segments took the
  // FP32 sum sweep from 4.0 to 3.4 ms, 16 -> 32 target-major segments the
  // gradient sweep from 3.4 to 2.9 ms); the q-major PASS 1 reads every
  // segment's maximum, which caps K1.  Large grids still gain from shorter
  // items (5.4M q-points / 500,000 targets: 4 -> 16 q-major segments took the
  // FP32 sum sweep from 26.9 to 21.6 ms, 4 -> 64 target-major segments the
  // gradient sweep from 50.8 to 24.1 ms), hence the floors of 16 and 64; the
  // segment partials (3 K1 nQ and K2 N doubles) stay below 2^28 entries per
  // array
  int64_t K1 = std::min<int64_t>(32, (65536 + p.n_blocks - 1) / std::max<int64_t>(1, p.n_blocks));
  int64_t K2 = (24576 + p.n_twaves - 1) / std::max<int64_t>(1, p.n_twaves);
  K1 = std::max<int64_t>(16, K1);
  K1 = std::max<int64_t>(1, std::min<int64_t>(K1, (1ll << 28) / std::max<int64_t>(1, p.nQ)));
  K2 = std::max<int64_t>(64, K2);
  K2 = std::max<int64_t>(1, std::min<int64_t>(K2, (1ll << 28) / std::max<int64_t>(1, p.N)));
  if (p.dense) {
    K1 = std::min<int64_t>(K1, std::max<int64_t>(1, p.N / 32));
    K2 = std::min<int64_t>(K2, std::max<int64_t>(1, p.n_cells / 8));
  }
  p.K1 = (int)std::min<int64_t>(K1, kMaxSeg);
  p.K2 = (int)std::min<int64_t>(K2, kMaxSeg);
  // the gradient grid gives each XCD the segments s = xcd (mod 8)
  if (p.K2 >= 8) p.K2 -= p.K2 % 8;
  if (c->force_K1 > 0) p.K1 = std::min(c->force_K1, kMaxSeg);
  if (c->force_K2 > 0) p.K2 = std::min(c->force_K2, kMaxSeg);
  p.nFpart = (p.nQ + 255) / 256;
  p.nPpart = (p.N + 255) / 256;
  // The FP32 seed reads the previous evaluation's segment maxima: the same
  // work items (grouping, q-major segments, dense or truncated).  Its
  // candidates may differ with the radius: a larger one adds targets the seed
  // does not bound (the sweep re-sweeps an item whose maximum rose above its
  // seed), a smaller one removes some (re-swept when its maximum fell too far
  // below; a point that had no candidates in the item has the seed -inf,
  // below any new maximum).  The value bound takes the previous maxima as
  // lower bounds of the new ones, which holds while every previous candidate
  // remains one: bounded sweeps are seeded at the same or a larger radius
  // (and not in items with a point that had no candidate before).
  const bool same_items = c->hist_G == p.G && c->hist_small == p.small && c->hist_K1 == p.K1 &&
                          c->hist_dense == p.dense;
  const bool radius_ok = R == c->hist_R || !(p.prune_bits > 0) || R > c->hist_R;
  p.hist = (p.mode == kF32Lse && c->hist_valid && same_items && radius_ok) ? 1 : 0;
  p.hist_same = (p.hist && R == c->hist_R) ? 1 : 0;
  return SDOT_OK;
}

// Note that an FP32 evaluation of plan p at radius R left its potentials and
// segment maxima behind (the next evaluation's seed).
void note_history(sdot_ctx *c, double R) {
  const Plan &p = c->plan;
  c->hist_valid = true;
  c->hist_R = R;
  c->hist_G = p.G;
  c->hist_small = p.small;
  c->hist_K1 = p.K1;
  c->hist_dense = p.dense;
}

// The plan at radius R with every per-evaluation buffer of the context large
// enough for it.
int prepare(sdot_ctx *c, double R) {
  int rc = plan_eval(c, R);
  if (rc) return rc;
  Plan &p = c->plan;
  Dev &d = c->dev;
  const size_t needM = (size_t)p.K1 * p.nQ, needG = (size_t)p.K2 * p.N;
  if (c->cap_M < needM) {
    HIP_TRY(dalloc(d.Mpart, needM));
    HIP_TRY(dalloc(d.Spart, needM));
    HIP_TRY(dalloc(d.Tpart, needM));
    c->cap_M = needM;
    c->hist_valid = false;  // the seed reads the previous segment maxima
    p.hist = 0;
  }
  if (c->cap_G < needG) {
    HIP_TRY(dalloc(d.Gpart, needG));
    c->cap_G = needG;
  }
  if (c->cap_trec < (size_t)p.N * kRec) {
    HIP_TRY(dalloc(d.trec_r, (size_t)p.N * kRec));
    HIP_TRY(dalloc(d.trec_m, (size_t)p.N * kRec));
    c->cap_trec = (size_t)p.N * kRec;
  }
  if (c->cap_qrec < (size_t)p.nQ * kRec) {
    HIP_TRY(dalloc(d.qrec_r, (size_t)p.nQ * kRec));
    HIP_TRY(dalloc(d.qchunk, (size_t)p.nQ * (p.dim + 2)));
    HIP_TRY(dalloc(d.Mglob, (size_t)p.nQ));
    c->cap_qrec = (size_t)p.nQ * kRec;
  }
  if (c->cap_F < (size_t)p.nFpart) {
    HIP_TRY(dalloc(d.Fpart, p.nFpart));
    HIP_TRY(dalloc(d.Cpart, p.nFpart));
    HIP_TRY(dalloc(d.Upart, p.nFpart));
    c->cap_F = p.nFpart;
  }
  if (c->cap_P < (size_t)p.nPpart) {
    HIP_TRY(dalloc(d.PNpart, p.nPpart));
    c->cap_P = p.nPpart;
  }
  if (c->cap_N < (size_t)p.N) {
    HIP_TRY(dalloc(d.psi, p.N));
    HIP_TRY(dalloc(d.grad, p.N));
    HIP_TRY(dalloc(d.psi_prev, p.N));
    HIP_TRY(dalloc(d.dpsi_r, p.N));
    HIP_TRY(dalloc(d.dpsi_cx, (p.N + 63) / 64));
    HIP_TRY(dalloc(d.dpsi_cn, (p.N + 63) / 64));
    c->cap_N = p.N;
    c->hist_valid = false;
    p.hist = 0;
  }
  if (!d.out) HIP_TRY(dalloc(d.out, 4));
  if (!d.ucount) HIP_TRY(dalloc(d.ucount, 1));
  if (p.prune_bits > 0) {
    const size_t nt = (size_t)p.N, ntb = (size_t)p.n_tbins, ntr = (size_t)grid_rows(p.tg);
    const size_t nc = (size_t)p.n_cells, ncb = (size_t)p.n_cbins, ncr = (size_t)grid_rows(p.cg);
    if (c->cap_vb_t < nt + ntb + ntr) {
      HIP_TRY(dalloc(d.psit_r, nt));
      HIP_TRY(dalloc(d.pbin, ntb));
      HIP_TRY(dalloc(d.prow, ntr));
      c->cap_vb_t = nt + ntb + ntr;
    }
    if (c->cap_vb_c < nc + ncb + ncr) {
      HIP_TRY(dalloc(d.vcell, nc));
      HIP_TRY(dalloc(d.vbin, ncb));
      HIP_TRY(dalloc(d.vrow, ncr));
      c->cap_vb_c = nc + ncb + ncr;
    }
  }
  return SDOT_OK;
}

int run_eval(sdot_ctx *c, double R, const double *d_psi, double *d_grad, double *d_FC) {
  int rc = prepare(c, R);
  if (rc) return rc;
  Profile pr;
  Profile *pp = nullptr;
  if (c->prof_on) {
    if (c->ev_pending == kProfSets) {
      rc = resolve_profile(c);
      if (rc) return rc;
    }
    pr.on = true;
    for (int k = 0; k < 6; ++k) pr.ev[k] = c->ev_pool[c->ev_pending][k];
    c->ev_pending++;
    pp = &pr;
  }
  HIP_TRY(launch_eval(c->plan, c->dev, d_psi, d_grad, d_FC, c->stream, pp));
  // an FP32 evaluation leaves psi and M_q behind to seed the next one
  if (c->plan.mode == kF32Lse) note_history(c, R);
  return SDOT_OK;
}

// The per-evaluation buffers of plan p laid out in one block at `base`
// (sub-arrays on 256-byte boundaries): d takes the context's source and target
// arrays from `ctx` and the evaluation's own arrays from the block.  Returns
// the block's size in bytes; base == nullptr sizes it only.
size_t carve_eval(const Plan &p, const Dev &ctx, double *base, Dev &d) {
  size_t off = 0;  // doubles
  auto take = [&](double *&dst, size_t n) {
    dst = base ? base + off : nullptr;
    off += (std::max<size_t>(n, 1) + 31) / 32 * 32;
  };
  const size_t N = (size_t)p.N, nQ = (size_t)p.nQ, M = (size_t)p.K1 * nQ;
  d = ctx;
  take(d.trec_r, N * kRec);
  take(d.trec_m, N * kRec);
  take(d.qrec_r, nQ * kRec);
  take(d.qchunk, nQ * (size_t)(p.dim + 2));
  take(d.Mpart, M);
  take(d.Spart, M);
  take(d.Tpart, M);
  take(d.Mglob, nQ);
  take(d.Gpart, (size_t)p.K2 * N);
  take(d.Fpart, (size_t)p.nFpart);
  take(d.Cpart, (size_t)p.nFpart);
  take(d.Upart, (size_t)p.nFpart);
  take(d.PNpart, (size_t)p.nPpart);
  take(d.psi_prev, N);
  take(d.dpsi_r, N);
  take(d.dpsi_cx, (N + 63) / 64);
  take(d.dpsi_cn, (N + 63) / 64);
  take(d.out, 4);
  take(d.ucount, 1);
  if (p.prune_bits > 0) {
    take(d.psit_r, N);
    take(d.pbin, (size_t)p.n_tbins);
    take(d.prow, (size_t)grid_rows(p.tg));
    take(d.vcell, (size_t)p.n_cells);
    take(d.vbin, (size_t)p.n_cbins);
    take(d.vrow, (size_t)grid_rows(p.cg));
  }
  return off * sizeof(double);
}

// sdot_eval_device without device-memory allocation: the evaluation's buffers
// are carved from this device's scratch array.  The FP32 sweep runs unseeded
// (the scratch keeps no history that belongs to this context), and the segment
// counts halve, the larger partial array first, until the buffers fit.
int run_eval_noalloc(sdot_ctx *c, double R, const double *d_psi, double *d_grad, double *d_FC) {
  if (!c->noalloc) return fail(SDOT_ERR_STATE, "no-allocation scratch is not set up on this device");
  int rc = plan_eval(c, R);
  if (rc) return rc;
  Plan p = c->plan;
  p.hist = 0;
  Dev d;
  size_t need = carve_eval(p, c->dev, nullptr, d);
  while (need > kNoAllocBytes && (p.K1 > 1 || p.K2 > 1)) {
    const size_t m = 3 * (size_t)p.K1 * (size_t)p.nQ, g = (size_t)p.K2 * (size_t)p.N;
    if (p.K1 > 1 && (m >= g || p.K2 == 1))
      p.K1 = (p.K1 + 1) / 2;
    else
      p.K2 = (p.K2 + 1) / 2;
    need = carve_eval(p, c->dev, nullptr, d);
  }
  if (need > kNoAllocBytes)
    return fail(SDOT_ERR_INVALID,
                "evaluation needs " + std::to_string(need) + " bytes of scratch with one segment per "
                "sweep; the no-allocation entry point holds " + std::to_string(kNoAllocBytes));
  carve_eval(p, c->dev, c->noalloc, d);
  HIP_TRY(launch_eval(p, d, d_psi, d_grad, d_FC, c->stream, nullptr));
  return SDOT_OK;
}

int host_eval(sdot_ctx *c, double R, const double *psi, double *grad, double *F, double *C) {
  if (!psi) return fail(SDOT_ERR_INVALID, "psi is null");
  const int64_t N = c->plan.N;
  if (!c->have_source || !c->have_target)
    return fail(SDOT_ERR_STATE, "Invalid measures configuration: source and target must be set");
  int rc = prepare(c, R);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(c->dev.psi, psi, sizeof(double) * N, hipMemcpyHostToDevice, c->stream));
  const bool seeded = c->plan.hist != 0;
  rc = run_eval(c, R, c->dev.psi, c->dev.grad, c->dev.out);
  if (rc) return rc;
  if (seeded) {
    // a seed far above the new maxima would have flushed the FP32 terms: redo unseeded
    double lost = 0.0;
    HIP_TRY(hipMemcpyAsync(&lost, c->dev.ucount, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (lost > 0.0) {
      c->hist_valid = false;
      HIP_TRY(hipMemcpyAsync(c->dev.psi, psi, sizeof(double) * N, hipMemcpyHostToDevice, c->stream));
      rc = prepare(c, R);
      if (rc) return rc;
      rc = run_eval(c, R, c->dev.psi, c->dev.grad, c->dev.out);
      if (rc) return rc;
    }
  }
  double fc[2];
  if (grad)
    HIP_TRY(hipMemcpyAsync(grad, c->dev.grad, sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipMemcpyAsync(fc, c->dev.out, sizeof(double) * 2, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (c->prof_on) {
    rc = resolve_profile(c);
    if (rc) return rc;
  }
  if (F) *F = fc[0];
  if (C) *C = fc[1];
  return SDOT_OK;
}

// The evaluation at R with the plan's first moments per target (barycenters,
// caller order, host memory).
int host_barycenters(sdot_ctx *c, double R, const double *psi, double *bary) {
  if (!psi || !bary) return fail(SDOT_ERR_INVALID, "null argument");
  if (!c->have_source || !c->have_target)
    return fail(SDOT_ERR_STATE, "Invalid measures configuration: source and target must be set");
  const int64_t N = c->plan.N;
  const int dim = c->plan.dim;
  int rc = prepare(c, R);
  if (rc) return rc;
  double *d_bary = nullptr, *d_B = nullptr;
  auto cleanup = [&]() {
    if (d_bary) (void)hipFree(d_bary);
    if (d_B) (void)hipFree(d_B);
  };
  hipError_t e = hipMalloc((void **)&d_bary, sizeof(double) * N * dim);
  if (e == hipSuccess) e = hipMalloc((void **)&d_B, sizeof(double) * (size_t)c->plan.K2 * dim * N);
  if (e == hipSuccess)
    e = hipMemcpyAsync(c->dev.psi, psi, sizeof(double) * N, hipMemcpyHostToDevice, c->stream);
  for (int attempt = 0; e == hipSuccess && attempt < 2; ++attempt) {
    const bool seeded = c->plan.hist != 0;
    e = launch_eval_bary(c->plan, c->dev, c->dev.psi, c->dev.grad, c->dev.out, d_bary, d_B,
                         c->stream);
    if (e != hipSuccess) break;
    if (c->plan.mode == kF32Lse) note_history(c, R);
    if (!seeded) break;
    // a seed far above the new maxima would have flushed the FP32 terms: redo unseeded
    double lost = 0.0;
    e = hipMemcpyAsync(&lost, c->dev.ucount, sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess || !(lost > 0.0)) break;
    c->hist_valid = false;
    rc = prepare(c, R);
    if (rc) {
      cleanup();
      return rc;
    }
  }
  if (e == hipSuccess)
    e = hipMemcpyAsync(bary, d_bary, sizeof(double) * N * dim, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  cleanup();
  if (e != hipSuccess) return fail(SDOT_ERR_HIP, std::string("barycenters: ") + hipGetErrorString(e));
  return SDOT_OK;
}

int raw_covering_radius(sdot_ctx *c, double *r0) {
  // a function of the cell centers and the targets only: kept until either changes
  if (c->cover_r0 >= 0.0) {
    *r0 = c->cover_r0;
    return SDOT_OK;
  }
  const Plan &p = c->plan;
  double *d_min = nullptr;
  HIP_TRY(hipMalloc((void **)&d_min, sizeof(double) * std::max<int64_t>(1, p.n_cells)));
  hipError_t e = launch_covering(p, c->dev, d_min, c->stream);
  std::vector<double> h(p.n_cells);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h.data(), d_min, sizeof(double) * p.n_cells, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d_min);
  if (e != hipSuccess) return fail(SDOT_ERR_HIP, std::string("covering radius: ") + hipGetErrorString(e));
  double m = 0.0;
  for (double v : h) m = std::max(m, v);
  *r0 = std::sqrt(m);
  c->cover_r0 = *r0;
  return SDOT_OK;
}

double vmax(const double *v, int64_t n) {
  double m = v[0];
  for (int64_t i = 1; i < n; ++i) m = v[i] > m ? v[i] : m;
  return m;
}
double vmin(const double *v, int64_t n) {
  double m = v[0];
  for (int64_t i = 1; i < n; ++i) m = v[i] < m ? v[i] : m;
  return m;
}

// compute_distance_threshold (SotSolver.cc:379-429) with its three bounds:
// integral (SotSolver.cc:475-507), geometric (SotSolver.cc:509-551) and
// pointwise (SotSolver.cc:406-419).
// The radius of the stored threshold rule at a potential with maximum mx and
// minimum mn (mn is read by the geometric rule only).
int threshold_from(sdot_ctx *c, double mx, double mn, double *R) {
  const int64_t N = c->plan.N;
  const double eps = c->params.epsilon, tau = c->params.tau;
  const int type = c->params.distance_threshold_type;
  if (type == SDOT_THRESH_INTEGRAL) {
    double absF = std::fabs(c->global_functional);
    if (absF < 1e-10) absF = 1e-10;
    const double r2 = 2.0 * mx + 2.0 * eps * std::log((eps * c->C_global) / (tau * absF));
    double r = std::sqrt(0.0 < r2 ? r2 : 0.0);
    *R = r < c->covering_radius ? c->covering_radius : r;
  } else if (type == SDOT_THRESH_GEOMETRIC) {
    double r0 = c->covering_radius;
    if (!(r0 > 0.0)) {
      int rc = raw_covering_radius(c, &r0);
      if (rc) return rc;
    }
    const double range = mx - mn;
    const double mind = c->min_target_density > 0.0 ? c->min_target_density : vmin(c->h_nu.data(), N);
    double fv = std::fabs(c->global_functional);
    if (fv < 1e-10) fv = 1e-10;
    const double lt = std::log(eps / (mind * tau * fv));
    double r2 = r0 * r0 + 2.0 * range + 2.0 * eps * lt;
    if (r2 < r0 * r0) r2 = r0 * r0;
    *R = std::sqrt(r2);
  } else {
    const double mind = c->min_target_density > 0.0 ? c->min_target_density : vmin(c->h_nu.data(), N);
    if (mind <= 0 || tau <= 0) {
      *R = DBL_MAX;
    } else {
      const double r2 = -2.0 * eps * std::log(tau / mind) + 2.0 * mx;
      *R = std::sqrt(0.0 < r2 ? r2 : 0.0);
    }
  }
  return SDOT_OK;
}

int threshold(sdot_ctx *c, const double *psi, double *R) {
  const int64_t N = c->plan.N;
  const bool geometric = c->params.distance_threshold_type == SDOT_THRESH_GEOMETRIC;
  return threshold_from(c, vmax(psi, N), geometric ? vmin(psi, N) : 0.0, R);
}

// evaluate_functional: threshold from the stored state, assembly, state update.
int stateful_eval(sdot_ctx *c, const double *psi, double *grad, double *F, double *C) {
  double R;
  int rc = threshold(c, psi, &R);
  if (rc) return rc;
  c->current_threshold = R;
  double f = 0, cg = 0;
  rc = host_eval(c, R, psi, grad, &f, &cg);
  if (rc) return rc;
  c->global_functional = f;
  c->C_global = cg;
  if (F) *F = f;
  if (C) *C = cg;
  return SDOT_OK;
}

// The statistics of launch_lb_stats for device vector a, reduced on the host
// in block order; d_part: kLbStats * kLbParts doubles on the device, h_part
// as many on the host.
hipError_t lb_stats(hipStream_t s, const double *a, const double *nu, double t, int64_t n,
                    double *d_part, double *h_part, double out[kLbStats]) {
  hipError_t e = launch_lb_stats(a, nu, t, n, d_part, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h_part, d_part, sizeof(double) * kLbStats * kLbParts, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  for (int q = 0; q < kLbStats; ++q) {
    const double *v = h_part + q * kLbParts;
    double r = v[0];
    for (int b = 1; b < kLbParts; ++b) {
      if (q == kLbL1 || q == kLbL2)
        r += v[b];
      else if (q == kLbMin)
        r = std::min(r, v[b]);
      else
        r = std::max(r, v[b]);
    }
    out[q] = r;
  }
  return hipSuccess;
}

// stateful_eval on a potential and gradient in HBM (caller order); the
// threshold rule reads the potential's maximum and minimum from a device
// reduction (d_part, h_part as in lb_stats).
int stateful_eval_device(sdot_ctx *c, const double *d_psi, double *d_grad, double *d_part,
                         double *h_part, double *F) {
  const int64_t N = c->plan.N;
  double st[kLbStats];
  HIP_TRY(lb_stats(c->stream, d_psi, nullptr, 0.0, N, d_part, h_part, st));
  double R;
  int rc = threshold_from(c, st[kLbMax], st[kLbMin], &R);
  if (rc) return rc;
  c->current_threshold = R;
  rc = prepare(c, R);
  if (rc) return rc;
  const bool seeded = c->plan.hist != 0;
  rc = run_eval(c, R, d_psi, d_grad, c->dev.out);
  if (rc) return rc;
  if (seeded) {
    // a seed far above the new maxima would have flushed the FP32 terms: redo unseeded
    double lost = 0.0;
    HIP_TRY(hipMemcpyAsync(&lost, c->dev.ucount, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (lost > 0.0) {
      c->hist_valid = false;
      rc = prepare(c, R);
      if (rc) return rc;
      rc = run_eval(c, R, d_psi, d_grad, c->dev.out);
      if (rc) return rc;
    }
  }
  double fc[2];
  HIP_TRY(hipMemcpyAsync(fc, c->dev.out, sizeof(double) * 2, hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (c->prof_on) {
    rc = resolve_profile(c);
    if (rc) return rc;
  }
  c->global_functional = fc[0];
  c->C_global = fc[1];
  *F = fc[0];
  return SDOT_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" const char *sdot_version(void) { return "sdot-mi355x 0.1 (gfx950)"; }

extern "C" const char *sdot_last_error(void) { return g_err.c_str(); }

extern "C" int sdot_ctx_create(int device, sdot_ctx **out) {
  if (!out) return fail(SDOT_ERR_INVALID, "out is null");
  *out = nullptr;
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(SDOT_ERR_INVALID, "no such HIP device");
  HIP_TRY(hipSetDevice(device));
  auto *c = new sdot_ctx();
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(SDOT_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  // the no-allocation entry point's scratch: its address on this device
  e = noalloc_scratch(device, &c->noalloc);
  if (e != hipSuccess) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return fail(SDOT_ERR_HIP, std::string("no-allocation scratch lookup: ") + hipGetErrorString(e));
  }
  sdot_params_default(&c->params);
  // test hook: SDOT_TEST_SEED_BIAS raises every seeded reference value of the
  // FP32 sweep by that many log2 units, which drives the sweep's re-sweep of
  // items whose seed lies too high (the seed bound itself never does)
  if (const char *b = std::getenv("SDOT_TEST_SEED_BIAS")) c->seed_bias = std::atof(b);
  // test hook: SDOT_TEST_SEGMENTS="K1,K2" fixes the segment counts of the
  // q-major and target-major sweeps (exercises counts the sizing rule picks
  // only at sizes no test reaches)
  if (const char *k = std::getenv("SDOT_TEST_SEGMENTS")) {
    int k1 = 0, k2 = 0;
    if (std::sscanf(k, "%d,%d", &k1, &k2) == 2 && k1 >= 1 && k2 >= 1) {
      c->force_K1 = k1;
      c->force_K2 = k2;
    }
  }
  *out = c;
  return SDOT_OK;
}

extern "C" int sdot_ctx_destroy(sdot_ctx *c) {
  if (!c) return SDOT_OK;
  if (c->n_servers > 0)
    return fail(SDOT_ERR_STATE, "a serving queue reads this context's source: destroy it first");
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  Dev &d = c->dev;
  void *bufs[] = {d.qx_m, d.qxr_m, d.qw_m, d.qx2_m, d.cc_m, d.blkbox, d.blkbox_s, d.qbox_s, d.cc_r, d.cstart, d.cell_r2m,
                  d.cell_m2o, d.ty_o, d.tnu_o, d.tpos_r, d.tpos_m, d.tr2o, d.tm2o, d.tstart,
                  d.twbox, d.ty_m, d.trec_r, d.trec_m, d.qrec_r, d.qchunk, d.Mpart, d.Spart, d.Gpart,
                  d.Fpart, d.Cpart, d.PNpart, d.psi, d.grad, d.out, d.scratch, d.i64scratch, d.Tpart,
                  d.Upart, d.psi_prev, d.dpsi_r, d.dpsi_cx, d.dpsi_cn, d.ucount, d.Mglob,
                  d.psit_r, d.pbin, d.prow, d.vcell, d.vbin, d.vrow};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (auto &set : c->ev_pool)
    for (auto e : set) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return SDOT_OK;
}

extern "C" int sdot_set_params(sdot_ctx *c, const sdot_params *p) {
  if (ctx_check(c) || !p) return fail(SDOT_ERR_INVALID, "null argument");
  if (!(p->epsilon > 0) || !std::isfinite(p->epsilon)) return fail(SDOT_ERR_INVALID, "epsilon must be finite and > 0");
  if (p->precision != SDOT_PREC_FP64 && p->precision != SDOT_PREC_FP32_LSE)
    return fail(SDOT_ERR_INVALID, "unknown precision");
  if (p->line_search != SDOT_LS_REFERENCE && p->line_search != SDOT_LS_NOISE_AWARE)
    return fail(SDOT_ERR_INVALID, "unknown line_search");
  if (p->preconditioner != SDOT_PRECOND_NONE && p->preconditioner != SDOT_PRECOND_DIAGONAL)
    return fail(SDOT_ERR_INVALID, "unknown preconditioner");
  if (std::isnan(p->prune_bits) || std::isinf(p->prune_bits))
    return fail(SDOT_ERR_INVALID, "prune_bits must be finite");
  c->params = *p;
  c->hist_valid = false;
  return SDOT_OK;
}

extern "C" int sdot_set_source(sdot_ctx *c, int dim, int64_t n_cells, int nq, const double *x,
                               const double *w, const double *centers) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  if (dim != 2 && dim != 3) return fail(SDOT_ERR_INVALID, "dim must be 2 or 3");
  if (n_cells < 1 || n_cells >= (1ll << 31) - 1) return fail(SDOT_ERR_INVALID, "n_cells out of range");
  if (nq < 1 || nq > 4096) return fail(SDOT_ERR_INVALID, "nq out of range");
  if (!x || !w || !centers) return fail(SDOT_ERR_INVALID, "null source array");
  if (c->n_servers > 0)
    return fail(SDOT_ERR_STATE, "a serving queue reads this context's source: destroy it first");
  if (c->have_target && dim != c->plan.dim)
    return fail(SDOT_ERR_INVALID, "source dim differs from the target dim");
  const int64_t nQ = n_cells * nq;
  if (!all_finite(x, nQ * dim) || !all_finite(w, nQ) || !all_finite(centers, n_cells * dim))
    return fail(SDOT_ERR_INVALID, "non-finite source values");
  HIP_TRY(hipSetDevice(c->device));
  c->cover_r0 = -1.0;
  Plan &p = c->plan;
  p.dim = dim;
  p.n_cells = n_cells;
  p.nq = nq;
  p.nQ = nQ;
  auto grouping = [&](int gmax) {
    Grouping gr;
    for (int g = gmax; g >= 1; --g)
      if (nq % g == 0) {
        gr.G = g;
        break;
      }
    gr.gpc = nq / gr.G;
    gr.n_groups = n_cells * gr.gpc;
    gr.n_blocks = (gr.n_groups + kWave - 1) / kWave;
    return gr;
  };
  p.grp_big = grouping(9);
  p.grp_small = grouping(4);
  p.small = 0;
  p.G = p.grp_big.G;
  p.gpc = p.grp_big.gpc;
  p.n_groups = p.grp_big.n_groups;
  p.n_blocks = p.grp_big.n_blocks;

  // bounding boxes; the shift origin is the middle of the quadrature-point box
  double qlo[3], qhi[3], clo[3], chi[3];
  for (int d = 0; d < dim; ++d) {
    qlo[d] = clo[d] = INFINITY;
    qhi[d] = chi[d] = -INFINITY;
  }
  for (int64_t i = 0; i < nQ; ++i)
    for (int d = 0; d < dim; ++d) {
      qlo[d] = std::min(qlo[d], x[i * dim + d]);
      qhi[d] = std::max(qhi[d], x[i * dim + d]);
    }
  for (int64_t i = 0; i < n_cells; ++i)
    for (int d = 0; d < dim; ++d) {
      clo[d] = std::min(clo[d], centers[i * dim + d]);
      chi[d] = std::max(chi[d], centers[i * dim + d]);
    }
  double scale = 0.0;
  for (int d = 0; d < 3; ++d) {
    p.origin[d] = d < dim ? 0.5 * (qlo[d] + qhi[d]) : 0.0;
    c->src_lo[d] = d < dim ? clo[d] : 0.0;
    c->src_hi[d] = d < dim ? chi[d] : 0.0;
    if (d < dim) scale = std::max({scale, std::fabs(clo[d]), std::fabs(chi[d])});
  }
  c->coord_scale = std::max(c->coord_scale, scale);

  // Morton order of the cells (q-major passes)
  std::vector<uint64_t> key(n_cells);
  for (int64_t i = 0; i < n_cells; ++i) key[i] = morton_key(centers + i * dim, clo, chi, dim);
  std::vector<int64_t> m2o(n_cells);
  std::iota(m2o.begin(), m2o.end(), 0);
  std::stable_sort(m2o.begin(), m2o.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
  std::vector<double> qx((size_t)nQ * dim), qxr((size_t)nQ * dim), qw(nQ), qx2(nQ),
      ccm((size_t)n_cells * dim);
  for (int64_t m = 0; m < n_cells; ++m) {
    const int64_t o = m2o[m];
    for (int d = 0; d < dim; ++d) ccm[m * dim + d] = centers[o * dim + d];
    for (int k = 0; k < nq; ++k) {
      double s2 = 0.0;
      for (int d = 0; d < dim; ++d) {
        const double v = x[(o * nq + k) * dim + d] - p.origin[d];
        qx[(m * nq + k) * dim + d] = v;
        qxr[(m * nq + k) * dim + d] = x[(o * nq + k) * dim + d];
        s2 += v * v;
      }
      qx2[m * nq + k] = s2;
      qw[m * nq + k] = w[o * nq + k];
    }
  }
  // AABB of the cell centers of each 64-group block, for both groupings
  auto boxes = [&](const Grouping &gr) {
    std::vector<double> box((size_t)gr.n_blocks * 2 * dim);
    for (int64_t b = 0; b < gr.n_blocks; ++b) {
      double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      const int64_t g0 = b * kWave, g1 = std::min<int64_t>(gr.n_groups, g0 + kWave);
      for (int64_t g = g0; g < g1; ++g) {
        const int64_t cm = g / gr.gpc;
        for (int d = 0; d < dim; ++d) {
          lo[d] = std::min(lo[d], ccm[cm * dim + d]);
          hi[d] = std::max(hi[d], ccm[cm * dim + d]);
        }
      }
      for (int d = 0; d < dim; ++d) {
        box[b * 2 * dim + d] = lo[d];
        box[b * 2 * dim + dim + d] = hi[d];
      }
    }
    return box;
  };
  const std::vector<double> box = boxes(p.grp_big), box_s = boxes(p.grp_small);
  // AABB of the quadrature points of each block of the small grouping (value
  // bound of the bounded sweeps, which run that grouping)
  std::vector<double> qbox_s((size_t)p.grp_small.n_blocks * 2 * dim);
  {
    const Grouping &gr = p.grp_small;
    for (int64_t b = 0; b < gr.n_blocks; ++b) {
      double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
      const int64_t g0 = b * kWave, g1 = std::min<int64_t>(gr.n_groups, g0 + kWave);
      for (int64_t g = g0; g < g1; ++g) {
        const int64_t cm = g / gr.gpc;
        const int64_t q0 = cm * nq + (g - cm * gr.gpc) * gr.G;
        for (int k = 0; k < gr.G; ++k)
          for (int d = 0; d < dim; ++d) {
            lo[d] = std::min(lo[d], qxr[(q0 + k) * dim + d]);
            hi[d] = std::max(hi[d], qxr[(q0 + k) * dim + d]);
          }
      }
      for (int d = 0; d < dim; ++d) {
        qbox_s[b * 2 * dim + d] = lo[d];
        qbox_s[b * 2 * dim + dim + d] = hi[d];
      }
    }
  }
  // largest distance of a cell's quadrature points from its center (value
  // bound of the target-major pass), with a margin for its rounding
  double rq2 = 0.0;
  for (int64_t i = 0; i < n_cells; ++i)
    for (int k = 0; k < nq; ++k) {
      double s2 = 0.0;
      for (int d = 0; d < dim; ++d) {
        const double t = x[(i * nq + k) * dim + d] - centers[i * dim + d];
        s2 += t * t;
      }
      rq2 = std::max(rq2, s2);
    }
  p.rq = std::sqrt(rq2) * (1.0 + 1e-9) + 1e-12 * scale;
  // row (bin) order of the cells (target-major pass)
  p.cg = make_grid(clo, chi, dim, n_cells, 8.0);
  const int64_t ncb = (int64_t)p.cg.n[0] * p.cg.n[1] * p.cg.n[2];
  p.n_cbins = ncb;
  std::vector<int> cbin(n_cells);
  for (int64_t m = 0; m < n_cells; ++m) cbin[m] = bin_linear(p.cg, &ccm[m * dim], dim);
  std::vector<int64_t> r2m(n_cells);
  std::iota(r2m.begin(), r2m.end(), 0);
  std::stable_sort(r2m.begin(), r2m.end(), [&](int64_t a, int64_t b) { return cbin[a] < cbin[b]; });
  std::vector<int32_t> cstart(ncb + 1, 0);
  for (int64_t m = 0; m < n_cells; ++m) cstart[cbin[m] + 1]++;
  for (int64_t b = 0; b < ncb; ++b) cstart[b + 1] += cstart[b];
  std::vector<double> ccr((size_t)n_cells * 4, 0.0);
  for (int64_t r = 0; r < n_cells; ++r)
    for (int d = 0; d < dim; ++d) ccr[r * 4 + d] = ccm[r2m[r] * dim + d];

  Dev &dv = c->dev;
  HIP_TRY(upload(dv.qx_m, qx));
  HIP_TRY(upload(dv.qxr_m, qxr));
  HIP_TRY(upload(dv.qw_m, qw));
  HIP_TRY(upload(dv.qx2_m, qx2));
  HIP_TRY(upload(dv.cc_m, ccm));
  HIP_TRY(upload(dv.blkbox, box));
  HIP_TRY(upload(dv.blkbox_s, box_s));
  HIP_TRY(upload(dv.qbox_s, qbox_s));
  HIP_TRY(upload(dv.cc_r, ccr));
  HIP_TRY(upload(dv.cstart, cstart));
  HIP_TRY(upload(dv.cell_r2m, r2m));
  HIP_TRY(upload(dv.cell_m2o, m2o));
  c->h_cell_r2o.resize(n_cells);
  for (int64_t r = 0; r < n_cells; ++r) c->h_cell_r2o[r] = m2o[r2m[r]];
  c->cap_M = c->cap_qrec = c->cap_F = c->cap_vb_c = 0;
  c->hist_valid = false;
  c->cover_r0 = -1.0;
  c->have_source = true;
  return SDOT_OK;
}

extern "C" int sdot_set_target(sdot_ctx *c, int64_t n, const double *y, const double *nu) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  if (n < 1 || n >= (1ll << 31) - 1) return fail(SDOT_ERR_INVALID, "Target measure not set up (n must be >= 1)");
  if (!y || !nu) return fail(SDOT_ERR_INVALID, "null target array");
  if (!c->have_source) return fail(SDOT_ERR_STATE, "call sdot_set_source before sdot_set_target");
  const int dim = c->plan.dim;
  if (!all_finite(y, n * dim) || !all_finite(nu, n)) return fail(SDOT_ERR_INVALID, "non-finite target values");
  HIP_TRY(hipSetDevice(c->device));
  Plan &p = c->plan;
  TargetShape S;
  target_shape(dim, n, y, S);
  p.N = n;
  p.n_twaves = S.n_twaves;
  p.tg = S.tg;
  for (int d = 0; d < 3; ++d) {
    c->tgt_lo[d] = S.lo[d];
    c->tgt_hi[d] = S.hi[d];
  }
  c->coord_scale = std::max(c->coord_scale, S.scale);
  c->n_tbins = S.n_bins;
  p.n_tbins = S.n_bins;
  p.nu_min = *std::min_element(nu, nu + n);
  if (!(p.nu_min > 0.0)) p.nu_min = 0.0;
  c->have_target = false;  // until every target array is in place
  c->cover_r0 = -1.0;
  // the orders, bin offsets and wave boxes are built on the device
  Dev &dv = c->dev;
  std::vector<double> yv(y, y + n * dim), nv(nu, nu + n);
  HIP_TRY(upload(dv.ty_o, yv));
  HIP_TRY(upload(dv.tnu_o, nv));
  HIP_TRY(dalloc(dv.tpos_r, (size_t)n));
  HIP_TRY(dalloc(dv.tpos_m, (size_t)n));
  HIP_TRY(dalloc(dv.tr2o, (size_t)n));
  HIP_TRY(dalloc(dv.tm2o, (size_t)n));
  HIP_TRY(dalloc(dv.tstart, (size_t)S.n_bins + 1));
  HIP_TRY(dalloc(dv.twbox, (size_t)S.n_twaves * 2 * dim));
  HIP_TRY(dalloc(dv.ty_m, (size_t)n * 4));
  const size_t sb = layout_scratch_bytes(S);
  if (sb == SIZE_MAX) return fail(SDOT_ERR_HIP, "target layout: no sort storage size");
  void *scratch = nullptr;
  HIP_TRY(hipMalloc(&scratch, sb));
  hipError_t e = device_target_layout(dim, S, dv.ty_o, dv.tpos_r, dv.tr2o, dv.tpos_m, dv.tm2o,
                                      dv.tstart, dv.twbox, dv.ty_m, scratch, sb, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(scratch);
  if (e != hipSuccess) return fail(SDOT_ERR_HIP, std::string("target layout: ") + hipGetErrorString(e));
  c->h_nu = nv;
  c->cap_G = c->cap_trec = c->cap_P = c->cap_N = c->cap_vb_t = 0;
  c->hist_valid = false;
  c->have_target = true;
  return SDOT_OK;
}

extern "C" int sdot_target_layout(sdot_ctx *c, int32_t *pos_r, int32_t *pos_m, int32_t *tstart,
                                  int64_t *n_bins, double *grid) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  if (!c->have_target) return fail(SDOT_ERR_STATE, "target measure not set");
  HIP_TRY(hipSetDevice(c->device));
  if (n_bins) *n_bins = c->n_tbins;
  if (grid) {
    const Grid &g = c->plan.tg;
    for (int d = 0; d < 3; ++d) {
      grid[d] = g.o[d];
      grid[4 + d] = (double)g.n[d];
    }
    grid[3] = g.h;
  }
  const size_t N = (size_t)c->plan.N;
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (pos_r) HIP_TRY(hipMemcpy(pos_r, c->dev.tpos_r, N * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (pos_m) HIP_TRY(hipMemcpy(pos_m, c->dev.tpos_m, N * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (tstart)
    HIP_TRY(hipMemcpy(tstart, c->dev.tstart, (size_t)(c->n_tbins + 1) * sizeof(int32_t),
                      hipMemcpyDeviceToHost));
  return SDOT_OK;
}

extern "C" int sdot_eval_R(sdot_ctx *c, double R, const double *psi, double *grad, double *F,
                           double *C) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  HIP_TRY(hipSetDevice(c->device));
  return host_eval(c, R, psi, grad, F, C);
}

extern "C" int sdot_eval(sdot_ctx *c, const double *psi, double *grad, double *F, double *C) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  if (!c->have_source || !c->have_target)
    return fail(SDOT_ERR_STATE, "Invalid measures configuration: source and target must be set");
  if (!psi) return fail(SDOT_ERR_INVALID, "psi is null");
  HIP_TRY(hipSetDevice(c->device));
  return stateful_eval(c, psi, grad, F, C);
}

extern "C" int sdot_barycenters_R(sdot_ctx *c, double R, const double *psi, double *bary) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  HIP_TRY(hipSetDevice(c->device));
  return host_barycenters(c, R, psi, bary);
}

extern "C" int sdot_barycenters(sdot_ctx *c, const double *psi, double *bary) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  if (!c->have_source || !c->have_target)
    return fail(SDOT_ERR_STATE, "Invalid measures configuration");
  if (!psi) return fail(SDOT_ERR_INVALID, "psi is null");
  HIP_TRY(hipSetDevice(c->device));
  // compute_weighted_barycenters_euclidean (SotSolver.cc:871-890): the threshold
  // from the stored state; the functional state is left as it is
  double R;
  int rc = threshold(c, psi, &R);
  if (rc) return rc;
  c->current_threshold = R;
  return host_barycenters(c, R, psi, bary);
}

extern "C" int sdot_eval_device(sdot_ctx *c, double R, const double *d_psi, double *d_grad,
                                double *d_FC) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  if (!d_psi || !d_grad || !d_FC) return fail(SDOT_ERR_INVALID, "null device pointer");
  HIP_TRY(hipSetDevice(c->device));
  return run_eval(c, R, d_psi, d_grad, d_FC);
}

extern "C" int sdot_eval_device_noalloc(sdot_ctx *c, double R, const double *d_psi,
                                        double *d_grad, double *d_FC) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  if (!d_psi || !d_grad || !d_FC) return fail(SDOT_ERR_INVALID, "null device pointer");
  HIP_TRY(hipSetDevice(c->device));
  return run_eval_noalloc(c, R, d_psi, d_grad, d_FC);
}

extern "C" int sdot_synchronize(sdot_ctx *c) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return resolve_profile(c);
}

extern "C" int sdot_hessian(sdot_ctx *c, const double *psi, double *H, int64_t ld) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  if (!c->have_source || !c->have_target)
    return fail(SDOT_ERR_STATE, "Invalid measures configuration");
  if (!psi || !H) return fail(SDOT_ERR_INVALID, "null argument");
  const int64_t N = c->plan.N;
  if (ld < N) return fail(SDOT_ERR_INVALID, "ld < N");
  // two N x N FP64 matrices on the device (69 GB at the limit) and one on the host
  if (N > 65536) return fail(SDOT_ERR_UNSUPPORTED, "dense Hessian limited to N <= 65536");
  HIP_TRY(hipSetDevice(c->device));
  double R;
  int rc = threshold(c, psi, &R);
  if (rc) return rc;
  // p_qj in FP64 with the max shift, whatever the evaluation mode
  c->force_mode = kF64Lse;
  double F, C;
  rc = host_eval(c, R, psi, nullptr, &F, &C);
  c->force_mode = -1;
  if (rc) return rc;
  const Plan &p = c->plan;
  // the SYRK kernel takes 128-column blocks and stages of 16 quadrature rows
  const int64_t Npad = (N + 127) / 128 * 128;
  const int64_t budget = (int64_t)1 << 30;  // bytes of one P chunk
  int64_t rows = std::max<int64_t>(16, budget / (8 * Npad) / 16 * 16);
  rows = std::min<int64_t>(rows, (p.nQ + 15) / 16 * 16);
  double *dA = nullptr, *dP = nullptr, *dH = nullptr;
  auto cleanup = [&]() {
    if (dA) (void)hipFree(dA);
    if (dP) (void)hipFree(dP);
    if (dH) (void)hipFree(dH);
  };
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipError_t e = hipMalloc((void **)&dA, sizeof(double) * Npad * Npad);
  if (e == hipSuccess) e = hipMalloc((void **)&dP, sizeof(double) * rows * Npad);
  if (e == hipSuccess) e = hipMalloc((void **)&dH, sizeof(double) * N * N);
  if (e == hipSuccess) e = hipMemsetAsync(dA, 0, sizeof(double) * Npad * Npad, c->stream);
  if (e == hipSuccess && c->prof_on) {
    e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    if (e == hipSuccess) e = hipEventRecord(ev[0], c->stream);
  }
  for (int64_t q0 = 0; e == hipSuccess && q0 < p.nQ; q0 += rows)
    e = launch_hessian_chunk(p, c->dev, q0, rows, dP, dA, Npad, c->stream);
  if (e == hipSuccess && ev[1]) e = hipEventRecord(ev[1], c->stream);
  if (e == hipSuccess) e = launch_hessian_final(p, dA, Npad, c->dev.grad, c->dev.tnu_o, dH, c->stream);
  if (e == hipSuccess)
    e = hipMemcpy2DAsync(H, sizeof(double) * ld, dH, sizeof(double) * N, sizeof(double) * N, N,
                         hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess && ev[1]) {
    float ms = 0.0f;
    e = hipEventElapsedTime(&ms, ev[0], ev[1]);
    c->prof_acc.ms_hessian += ms;
    c->prof_acc.hessians += 1;
  }
  for (hipEvent_t x : ev)
    if (x) (void)hipEventDestroy(x);
  cleanup();
  if (e != hipSuccess) return fail(SDOT_ERR_HIP, std::string("hessian: ") + hipGetErrorString(e));
  return SDOT_OK;
}

extern "C" int sdot_softmax_refinement(sdot_ctx *c, int64_t n_fine, const double *y_fine,
                                       int64_t n_coarse, const double *y_coarse,
                                       const double *nu_coarse, const double *psi_coarse,
                                       int64_t n_parents, const int64_t *child_offsets,
                                       const int64_t *children, double lambda,
                                       double distance_threshold, int use_log_sum_exp_trick,
                                       double *psi_fine) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  if (!c->have_source) return fail(SDOT_ERR_STATE, "call sdot_set_source before the refinement");
  if (n_fine < 0 || n_fine >= (1ll << 31) - 1) return fail(SDOT_ERR_INVALID, "n_fine out of range");
  if (n_coarse < 1 || n_coarse >= (1ll << 31) - 1)
    return fail(SDOT_ERR_INVALID, "n_coarse out of range");
  if (n_parents < 0) return fail(SDOT_ERR_INVALID, "n_parents < 0");
  if ((n_fine > 0 && (!y_fine || !psi_fine)) || !y_coarse || !nu_coarse || !psi_coarse ||
      !child_offsets)
    return fail(SDOT_ERR_INVALID, "null argument");
  if (!(lambda > 0) || !std::isfinite(lambda))
    return fail(SDOT_ERR_INVALID, "regularization parameter must be finite and > 0");
  const int dim = c->plan.dim;
  if (!all_finite(y_coarse, n_coarse * dim) || !all_finite(nu_coarse, n_coarse) ||
      !all_finite(psi_coarse, n_coarse) || (n_fine > 0 && !all_finite(y_fine, n_fine * dim)))
    return fail(SDOT_ERR_INVALID, "non-finite coarse or fine values");
  if (child_offsets[0] != 0) return fail(SDOT_ERR_INVALID, "child_offsets[0] must be 0");
  for (int64_t i = 0; i < n_parents; ++i)
    if (child_offsets[i + 1] < child_offsets[i])
      return fail(SDOT_ERR_INVALID, "child_offsets must be non-decreasing");
  if (child_offsets[n_parents] > 0 && !children) return fail(SDOT_ERR_INVALID, "children is null");
  HIP_TRY(hipSetDevice(c->device));
  const Plan &p = c->plan;

  // coarse points in bin-row order (the q-major pass streams bin runs)
  TargetLayout L;
  build_target_layout(dim, n_coarse, y_coarse, L);
  std::vector<double> crec((size_t)n_coarse * kRec, 0.0);
  for (int64_t r = 0; r < n_coarse; ++r) {
    const int64_t j = L.r2o[r];
    for (int d = 0; d < dim; ++d) crec[r * kRec + d] = y_coarse[j * dim + d];
    crec[r * kRec + 3] = psi_coarse[j];
    crec[r * kRec + 4] = nu_coarse[j];
  }
  // (parent, child) entries of the child lists; coarse indices without a list
  // and child indices outside the fine level are skipped, as the reference
  // does after reporting them (SoftmaxRefinement.cc:109-128)
  std::vector<int64_t> ent_c, ent_f;
  for (int64_t ci = 0; ci < std::min(n_coarse, n_parents); ++ci)
    for (int64_t k = child_offsets[ci]; k < child_offsets[ci + 1]; ++k) {
      const int64_t f = children[k];
      if (f < 0 || f >= n_fine) continue;
      ent_c.push_back(ci);
      ent_f.push_back(f);
    }
  const int64_t E = (int64_t)ent_c.size();
  std::vector<int64_t> order(E);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    return L.pos_m[ent_c[a]] < L.pos_m[ent_c[b]];
  });
  std::vector<double> erec((size_t)std::max<int64_t>(E, 1) * kRec, 0.0);
  std::vector<int64_t> foff(n_fine + 1, 0), fent(std::max<int64_t>(E, 1), 0);
  for (int64_t k = 0; k < E; ++k) {
    const int64_t e = order[k];
    for (int d = 0; d < dim; ++d) {
      erec[k * kRec + d] = y_coarse[ent_c[e] * dim + d];
      erec[k * kRec + 3 + d] = y_fine[ent_f[e] * dim + d];
    }
    foff[ent_f[e] + 1]++;
  }
  for (int64_t f = 0; f < n_fine; ++f) foff[f + 1] += foff[f];
  {
    std::vector<int64_t> fill(foff.begin(), foff.end() - 1);
    for (int64_t k = 0; k < E; ++k) fent[fill[ent_f[order[k]]]++] = k;
  }

  RefArgs a{};
  a.tg = L.shape.tg;
  a.cg = p.cg;
  a.nQ = p.nQ;
  a.n_cells = p.n_cells;
  a.Nc = n_coarse;
  a.E = E;
  a.n_fine = n_fine;
  a.nq = p.nq;
  a.lse = use_log_sum_exp_trick ? 1 : 0;
  a.smax = smax_for(distance_threshold);
  a.s2max = s2max_for(distance_threshold);
  const double scale = std::max(c->coord_scale, L.shape.scale);
  a.Rq = std::isnan(distance_threshold)
             ? 0.0
             : distance_threshold * (1.0 + 1e-9) + 1e-9 * (L.shape.tg.h + p.cg.h + scale);
  a.lam = lambda;
  a.lam_inv = 1.0 / lambda;
  const int64_t nwq = (p.nQ + kWave - 1) / kWave, nwe = (E + kWave - 1) / kWave;
  a.K1 = (int)std::min<int64_t>(64, std::max<int64_t>(1, 8192 / std::max<int64_t>(1, nwq)));
  a.K2 = (int)std::min<int64_t>(64, std::max<int64_t>(1, 16384 / std::max<int64_t>(1, nwe)));
  if (a.K2 >= 8) a.K2 -= a.K2 % 8;
  a.qxr = c->dev.qxr_m;
  a.qw = c->dev.qw_m;
  a.cc = c->dev.cc_m;
  a.cc_r = c->dev.cc_r;
  a.cstart = c->dev.cstart;
  a.cell_r2m = c->dev.cell_r2m;

  std::vector<void *> tmp;
  auto cleanup = [&]() {
    for (void *b : tmp) (void)hipFree(b);
  };
  auto alloc = [&](auto *&ptr, size_t n) {
    void *b = nullptr;
    hipError_t e = hipMalloc(&b, std::max<size_t>(n, 1) * sizeof(*ptr));
    if (e == hipSuccess) tmp.push_back(b);
    ptr = (std::remove_reference_t<decltype(ptr)>)b;
    return e;
  };
  double *d_crec = nullptr, *d_erec = nullptr, *d_M = nullptr, *d_S = nullptr, *d_rq = nullptr,
         *d_E = nullptr, *d_psi = nullptr;
  int32_t *d_tstart = nullptr;
  int64_t *d_foff = nullptr, *d_fent = nullptr;
  hipError_t e = alloc(d_crec, crec.size());
  if (e == hipSuccess) e = alloc(d_tstart, L.tstart.size());
  if (e == hipSuccess) e = alloc(d_erec, erec.size());
  if (e == hipSuccess) e = alloc(d_foff, foff.size());
  if (e == hipSuccess) e = alloc(d_fent, fent.size());
  if (e == hipSuccess) e = alloc(d_M, (size_t)a.K1 * p.nQ);
  if (e == hipSuccess) e = alloc(d_S, (size_t)a.K1 * p.nQ);
  if (e == hipSuccess) e = alloc(d_rq, (size_t)p.nQ * 4);
  if (e == hipSuccess) e = alloc(d_E, (size_t)a.K2 * std::max<int64_t>(E, 1));
  if (e == hipSuccess) e = alloc(d_psi, (size_t)n_fine);
  auto h2d = [&](void *dst, const void *src, size_t bytes) {
    if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream);
  };
  h2d(d_crec, crec.data(), crec.size() * sizeof(double));
  h2d(d_tstart, L.tstart.data(), L.tstart.size() * sizeof(int32_t));
  h2d(d_erec, erec.data(), erec.size() * sizeof(double));
  h2d(d_foff, foff.data(), foff.size() * sizeof(int64_t));
  h2d(d_fent, fent.data(), fent.size() * sizeof(int64_t));
  a.crec = d_crec;
  a.tstart = d_tstart;
  a.erec = d_erec;
  a.foff = d_foff;
  a.fent = d_fent;
  a.Mpart = d_M;
  a.Spart = d_S;
  a.rq = d_rq;
  a.Epart = d_E;
  a.psi_fine = d_psi;
  if (e == hipSuccess) e = launch_refinement(dim, a, c->stream);
  if (e == hipSuccess && n_fine > 0)
    e = hipMemcpyAsync(psi_fine, d_psi, sizeof(double) * n_fine, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  cleanup();
  if (e != hipSuccess) return fail(SDOT_ERR_HIP, std::string("softmax refinement: ") + hipGetErrorString(e));
  return SDOT_OK;
}

// Largest double s with sqrt(s) < R: the strict test of the transport maps'
// truncation (OptimalTransportPlan.cc:136, :262) as s <= smax.
static double smax_strict(double R) {
  if (std::isnan(R) || !(R > 0)) return -1.0;
  if (std::isinf(R)) return INFINITY;
  double s = R * R;
  if (std::isinf(s)) s = DBL_MAX;
  while (s > 0 && !(std::sqrt(s) < R)) s = std::nextafter(s, 0.0);
  for (;;) {
    const double n = std::nextafter(s, INFINITY);
    if (std::isinf(n) || !(std::sqrt(n) < R)) break;
    s = n;
  }
  return s;
}

// Point queries over the context's target measure (op 0: conditional
// densities at R; 1: barycentric map; 2: modal map, truncated at `radius`).
static int host_points(sdot_ctx *c, int op, const double *psi, double eps, double R,
                       int64_t n, const double *points, const double *density, int64_t n_idx,
                       const int64_t *indices, double *out) {
  if (!c->have_target) return fail(SDOT_ERR_STATE, "Invalid measures configuration: no target");
  if (n < 0 || n >= (1ll << 40)) return fail(SDOT_ERR_INVALID, "point count out of range");
  if (n_idx < 0) return fail(SDOT_ERR_INVALID, "index count < 0");
  // the density table has n * n_idx entries, one thread each
  if (n_idx > 0 && n > ((1ll << 40) - 1) / n_idx)
    return fail(SDOT_ERR_INVALID, "points x indices out of range");
  if (!psi || (n > 0 && (!points || !out)) || (op == 0 && n > 0 && (!density || (n_idx > 0 && !indices))))
    return fail(SDOT_ERR_INVALID, "null argument");
  if (!(eps > 0) || !std::isfinite(eps))
    return fail(SDOT_ERR_INVALID, "regularization parameter must be finite and > 0");
  if (std::isnan(R)) return fail(SDOT_ERR_INVALID, "radius is NaN");
  const Plan &p = c->plan;
  const int dim = p.dim;
  const int64_t N = p.N;
  if (!all_finite(psi, N) || !all_finite(points, n * dim))
    return fail(SDOT_ERR_INVALID, "non-finite potential or points");
  if (n == 0) return SDOT_OK;
  HIP_TRY(hipSetDevice(c->device));

  // query points in Morton order (waves of nearby points share their bin rows)
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int64_t i = 0; i < n; ++i)
    for (int d = 0; d < dim; ++d) {
      lo[d] = std::min(lo[d], points[i * dim + d]);
      hi[d] = std::max(hi[d], points[i * dim + d]);
    }
  std::vector<uint64_t> key(n);
  for (int64_t i = 0; i < n; ++i) key[i] = morton_key(points + i * dim, lo, hi, dim);
  std::vector<int64_t> m2o(n);
  std::iota(m2o.begin(), m2o.end(), 0);
  std::stable_sort(m2o.begin(), m2o.end(), [&](int64_t a, int64_t b) { return key[a] < key[b]; });
  std::vector<double> px((size_t)n * dim);
  for (int64_t m = 0; m < n; ++m)
    for (int d = 0; d < dim; ++d) px[m * dim + d] = points[m2o[m] * dim + d];
  std::vector<double> logterm;
  if (op == 2) {
    // the reference's log term, with the same libm log (OptimalTransportPlan.cc:170-175)
    const std::vector<double> &nu = c->h_nu;
    logterm.resize(N);
    for (int64_t j = 0; j < N; ++j) logterm[j] = nu[j] > 0 ? eps * std::log(nu[j]) : -INFINITY;
  }

  PtArgs a{};
  a.tg = p.tg;
  a.n = n;
  a.N = N;
  a.tstart = c->dev.tstart;
  a.tpos_r = c->dev.tpos_r;
  a.ty = c->dev.ty_o;
  a.nu = c->dev.tnu_o;
  a.eps = eps;
  a.eps_inv = 1.0 / eps;
  a.divide = op == 1 ? 1 : 0;
  a.moments = op == 1 ? 1 : 0;
  const bool trunc = op == 0 ? !(R >= DBL_MAX) : R > 0.0;
  a.truncated = op != 0 && trunc ? 1 : 0;
  a.smax = op == 0 ? smax_for(R) : (trunc ? smax_strict(R) : INFINITY);
  a.Rq = trunc ? R * (1.0 + 1e-9) + 1e-9 * (p.tg.h + c->coord_scale) : INFINITY;
  if (op == 0 && R < 0) a.Rq = 0.0;
  double span2 = 0.0;
  for (int d = 0; d < dim; ++d) {
    const double a0 = std::min(lo[d], p.tg.o[d]), a1 = std::max(hi[d], p.tg.o[d] + p.tg.n[d] * p.tg.h);
    span2 += (a1 - a0) * (a1 - a0);
  }
  a.rho_all = std::sqrt(span2) * (1.0 + 1e-9) + 1e-9;
  a.rho0 = std::min(a.rho_all, std::max(trunc ? R : 0.0, p.tg.h));
  const int64_t nw = (n + kWave - 1) / kWave;
  a.K = (int)std::min<int64_t>(64, std::max<int64_t>(1, 8192 / nw));
  if (a.K >= 8) a.K -= a.K % 8;
  a.n_idx = op == 0 ? n_idx : 0;

  std::vector<void *> tmp;
  auto alloc = [&](auto *&ptr, size_t cnt) {
    void *b = nullptr;
    hipError_t e = hipMalloc(&b, std::max<size_t>(cnt, 1) * sizeof(*ptr));
    if (e == hipSuccess) tmp.push_back(b);
    ptr = (std::remove_reference_t<decltype(ptr)>)b;
    return e;
  };
  double *d_psi = nullptr, *d_px = nullptr, *d_praw = nullptr, *d_rec = nullptr, *d_log = nullptr,
         *d_M = nullptr,
         *d_S = nullptr, *d_B = nullptr, *d_Mo = nullptr, *d_So = nullptr, *d_map = nullptr,
         *d_rho = nullptr, *d_cond = nullptr;
  int64_t *d_m2o = nullptr, *d_C = nullptr, *d_I = nullptr, *d_idx = nullptr;
  int32_t *d_need = nullptr;
  const size_t Kn = (size_t)a.K * n;
  hipError_t e = alloc(d_px, (size_t)n * dim);
  if (e == hipSuccess) e = alloc(d_psi, (size_t)N);
  if (e == hipSuccess) e = alloc(d_m2o, (size_t)n);
  if (e == hipSuccess) e = alloc(d_rec, (size_t)N * kRec);
  if (e == hipSuccess) e = alloc(d_M, Kn);
  if (e == hipSuccess) e = alloc(d_C, Kn);
  if (e == hipSuccess && op != 2) e = alloc(d_S, Kn);
  if (e == hipSuccess && op == 1) e = alloc(d_B, Kn * dim);
  if (e == hipSuccess && op == 2) e = alloc(d_I, Kn);
  if (e == hipSuccess && op == 2) e = alloc(d_log, (size_t)N);
  if (e == hipSuccess && op != 0) e = alloc(d_need, (size_t)n);
  if (e == hipSuccess && op != 0) e = alloc(d_map, (size_t)n * dim);
  if (e == hipSuccess && op == 0) e = alloc(d_Mo, (size_t)n);
  if (e == hipSuccess && op == 0) e = alloc(d_So, (size_t)n);
  if (e == hipSuccess && op == 0) e = alloc(d_praw, (size_t)n * dim);
  if (e == hipSuccess && op == 0) e = alloc(d_rho, (size_t)n);
  if (e == hipSuccess && op == 0) e = alloc(d_idx, (size_t)n_idx);
  if (e == hipSuccess && op == 0) e = alloc(d_cond, (size_t)n_idx * n);
  auto h2d = [&](void *dst, const void *src, size_t bytes) {
    if (e == hipSuccess && bytes > 0)
      e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream);
  };
  h2d(d_psi, psi, sizeof(double) * N);
  h2d(d_px, px.data(), px.size() * sizeof(double));
  h2d(d_m2o, m2o.data(), m2o.size() * sizeof(int64_t));
  if (op == 2) h2d(d_log, logterm.data(), logterm.size() * sizeof(double));
  if (op == 0) {
    h2d(d_praw, points, sizeof(double) * n * dim);
    h2d(d_rho, density, sizeof(double) * n);
    h2d(d_idx, indices, sizeof(int64_t) * n_idx);
  }
  a.px = d_px;
  a.psi = d_psi;
  a.praw = d_praw;
  a.pm2o = d_m2o;
  a.rec_w = d_rec;
  a.logterm = d_log;
  a.Mpart = d_M;
  a.Spart = d_S;
  a.Bpart = d_B;
  a.Cpart = d_C;
  a.Ipart = d_I;
  a.Mo = d_Mo;
  a.So = d_So;
  a.need = d_need;
  a.mapped = d_map;
  a.idx = d_idx;
  a.rho = d_rho;
  a.cond = d_cond;
  if (e == hipSuccess) e = launch_points(dim, a, op, c->stream);
  if (e == hipSuccess) {
    if (op == 0 && n_idx > 0)
      e = hipMemcpyAsync(out, d_cond, sizeof(double) * n_idx * n, hipMemcpyDeviceToHost, c->stream);
    else if (op != 0)
      e = hipMemcpyAsync(out, d_map, sizeof(double) * n * dim, hipMemcpyDeviceToHost, c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  for (void *b : tmp) (void)hipFree(b);
  if (e != hipSuccess) return fail(SDOT_ERR_HIP, std::string("point query: ") + hipGetErrorString(e));
  return SDOT_OK;
}

extern "C" int sdot_conditional_density(sdot_ctx *c, const double *psi, double R,
                                        int64_t n_points, const double *points,
                                        const double *density, int64_t n_idx,
                                        const int64_t *indices, double *out) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  return host_points(c, 0, psi, c->params.epsilon, R, n_points, points, density, n_idx, indices,
                     out);
}

extern "C" int sdot_transport_map(sdot_ctx *c, int strategy, const double *psi, double epsilon,
                                  double radius, int64_t n_points, const double *points,
                                  double *mapped) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  if (strategy != SDOT_MAP_MODAL && strategy != SDOT_MAP_BARYCENTRIC)
    return fail(SDOT_ERR_INVALID, "Unknown strategy");
  return host_points(c, strategy == SDOT_MAP_MODAL ? 2 : 1, psi, epsilon, radius, n_points,
                     points, nullptr, 0, nullptr, mapped);
}

// Potentials at least this long are solved with every L-BFGS vector in HBM
// (the iterate, gradient, direction and the 5 (s, y) pairs): on one host
// thread the two-loop recursion over 500,000-vectors alone takes ~10 ms, and
// x and g would cross PCIe at every evaluation.  Measured per evaluation of
// 3D Gaussian-mixture solves, host / device: 4,000 targets 0.30 / 0.39 ms,
// 16,000 1.40 / 1.44 ms, 50,000 6.15 / 5.88 ms (FP32).
// SDOT_LBFGS_DEVICE_MIN (a non-negative integer) overrides the length.
constexpr int64_t kDevicePairsMin = 32768;
static int64_t device_pairs_min() {
  const char *e = std::getenv("SDOT_LBFGS_DEVICE_MIN");
  if (e && *e) {
    char *end = nullptr;
    const long long v = std::strtoll(e, &end, 10);
    if (end && *end == 0 && v >= 0) return (int64_t)v;
  }
  return kDevicePairsMin;
}

// The space of lbfgs_core on device vectors (sdot_lbfgs.hip): the iterate
// x(), the work vectors, the pairs, diag(h0) and the target weights in one
// allocation.  Dot products are reduced on the host from kLbParts partial
// sums in block order.  The first HIP error ends the solve and is kept.
class DeviceSpace {
 public:
  hipError_t err = hipSuccess;
  const char *what = "";

  bool init(int64_t n, int history, const double *h0, const double *nu, hipStream_t s) {
    n_ = n;
    stream_ = s;
    history_ = history;
    if (history > kLbMaxHist) return check(hipErrorInvalidValue, "history length");
    const size_t vn = (size_t)n;
    const size_t words = (size_t)(kNumVecs + 1 + 2 * history + 1 + (h0 ? 1 : 0)) * vn +
                         (size_t)kLbStats * kLbParts + (size_t)history;
    if (!check(hipMalloc((void **)&mem_, words * sizeof(double)), "hipMalloc")) return false;
    if (!check(hipHostMalloc((void **)&hpart_, sizeof(double) * kLbStats * kLbParts,
                             hipHostMallocDefault), "hipHostMalloc"))
      return false;
    double *w = mem_;
    for (int i = 0; i < kNumVecs; ++i, w += vn) work_[i] = w;
    x_ = w;
    w += vn;
    for (int i = 0; i < history; ++i, w += vn) S_[i] = w;
    for (int i = 0; i < history; ++i, w += vn) Y_[i] = w;
    nu_ = w;
    w += vn;
    if (!upload(nu_, nu, "weights upload")) return false;
    if (h0) {
      h0_ = w;
      w += vn;
      if (!upload(h0_, h0, "h0 upload")) return false;
    }
    part_ = w;
    coef_ = w + (size_t)kLbStats * kLbParts;
    return true;
  }
  ~DeviceSpace() {
    if (mem_) {
      (void)hipStreamSynchronize(stream_);
      (void)hipFree(mem_);
    }
    if (hpart_) (void)hipHostFree(hpart_);
  }

  double *x() { return x_; }
  double *vec(int i) { return work_[i]; }
  double *part() { return part_; }
  double *hpart() { return hpart_; }
  const double *nu() const { return nu_; }
  bool ok() const { return err == hipSuccess; }
  bool check(hipError_t e, const char *where) {
    if (e != hipSuccess && err == hipSuccess) {
      err = e;
      what = where;
    }
    return e == hipSuccess;
  }
  bool upload(double *d, const double *h, const char *where) {
    return check(hipMemcpyAsync(d, h, sizeof(double) * n_, hipMemcpyHostToDevice, stream_), where);
  }
  bool download(double *h, const double *d, const char *where) {
    return check(hipMemcpyAsync(h, d, sizeof(double) * n_, hipMemcpyDeviceToHost, stream_), where) &&
           check(hipStreamSynchronize(stream_), where);
  }

  void copy(double *d, const double *s) {
    if (ok()) check(hipMemcpyAsync(d, s, sizeof(double) * n_, hipMemcpyDeviceToDevice, stream_), "copy");
  }
  double dot(const double *a, const double *b) {
    if (!ok() || !check(launch_lb_dot(a, b, n_, part_, stream_), "dot") ||
        !check(hipMemcpyAsync(hpart_, part_, sizeof(double) * kLbParts, hipMemcpyDeviceToHost,
                              stream_), "dot") ||
        !check(hipStreamSynchronize(stream_), "dot"))
      return NAN;
    double r = hpart_[0];
    for (int b = 1; b < kLbParts; ++b) r += hpart_[b];
    return r;
  }
  void lin(double *x, const double *x0, double a, const double *p) {
    if (ok()) check(launch_lb_lin(x, x0, a, p, n_, stream_), "x0 + a p");
  }
  void diff(double *sv, const double *x, double *y, const double *g) {
    if (ok()) check(launch_lb_diff(sv, x, y, g, n_, stream_), "pair differences");
  }
  void store(int slot) {
    std::swap(S_[slot], work_[kVecS]);
    std::swap(Y_[slot], work_[kVecY]);
  }
  void direction(const double *g, double *p, const int *slots, const double *rho, int count) {
    if (!ok()) return;
    LbVecs L{};
    for (int i = 0; i < history_; ++i) {
      L.S[i] = S_[i];
      L.Y[i] = Y_[i];
    }
    L.n = n_;
    L.p = p;
    L.h0 = h0_;
    L.part = part_;
    L.coef = coef_;
    copy(p, g);
    if (ok()) check(launch_lb_direction(L, slots, rho, count, stream_), "two-loop recursion");
  }
  // sum |a|, sum a^2, max (|a| - t nu) of a device vector (NaN after a failure)
  bool stats(const double *a, double t, double out[kLbStats]) {
    if (ok() && check(lb_stats(stream_, a, nu_, t, n_, part_, hpart_, out), "gradient statistics"))
      return true;
    for (int q = 0; q < kLbStats; ++q) out[q] = NAN;
    return false;
  }

 private:
  int64_t n_ = 0;
  int history_ = 0;
  hipStream_t stream_ = nullptr;
  double *mem_ = nullptr, *hpart_ = nullptr;
  double *work_[kNumVecs] = {}, *x_ = nullptr, *S_[kLbMaxHist] = {}, *Y_[kLbMaxHist] = {};
  double *nu_ = nullptr, *h0_ = nullptr, *part_ = nullptr, *coef_ = nullptr;
};

extern "C" int sdot_solve(sdot_ctx *c, double *psi, sdot_solve_info *info) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  if (!c->have_source || !c->have_target)
    return fail(SDOT_ERR_STATE, "Invalid measures configuration");
  if (!psi) return fail(SDOT_ERR_INVALID, "psi is null");
  HIP_TRY(hipSetDevice(c->device));
  const auto t0 = std::chrono::steady_clock::now();
  const int64_t N = c->plan.N;
  double r0 = 0.0;
  int rc = raw_covering_radius(c, &r0);
  if (rc) return rc;
  c->covering_radius = r0 * 1.1;
  c->min_target_density = vmin(c->h_nu.data(), N);
  int eval_rc = SDOT_OK;
  const bool comp = c->params.solver_control_type == SDOT_CTRL_COMPONENTWISE;
  const double tol = comp ? DBL_MIN : c->params.tolerance;
  const double utol = c->params.tolerance;
  const unsigned max_it = c->params.max_iterations;
  // the iteration line VerboseSolverControl::check prints (SotSolver.h:408-440)
  double l1_first = 0.0;
  auto report = [&](int step, double l2, double cv) {
    if (comp) {
      std::printf("Iteration %d - L-2 gradient norm: %g - Max Scaled Residual (max |g_i| - T_i*tol): %g\n",
                  step, l2, cv);
    } else {
      const double rel = (step == 0 || l1_first == 0.0) ? cv : cv / l1_first;
      if (step == 0) l1_first = cv;
      std::printf("Iteration %d - L-2 gradient norm: %g - L-1 gradient norm: %g - Relative L-1 residual: %g\n",
                  step, l2, cv, rel);
    }
    std::fflush(stdout);
  };
  const bool noise_aware = c->params.line_search == SDOT_LS_NOISE_AWARE;
  const bool low = c->params.use_log_sum_exp_trick && c->params.precision == SDOT_PREC_FP32_LSE;
  // initial inverse Hessian diag(eps / nu_j) (identity: the reference's)
  std::vector<double> h0v;
  if (c->params.preconditioner == SDOT_PRECOND_DIAGONAL) {
    // weights <= 0 (their targets receive no mass) take the mean weight's scale
    double mean = 0.0;
    for (int64_t j = 0; j < N; ++j) mean += std::fabs(c->h_nu[j]);
    mean = mean > 0.0 ? mean / (double)N : 1.0;
    h0v.resize(N);
    for (int64_t j = 0; j < N; ++j)
      h0v[j] = c->params.epsilon / (c->h_nu[j] > 0.0 ? c->h_nu[j] : mean);
  }
  const double *h0 = h0v.empty() ? nullptr : h0v.data();
  BfgsResult r;
  int its = 0, evals = 0;
  // The solve over a space's vectors: stat(g, mult, &l2) is the stopping value
  // at tolerance scale `mult` (SotSolver.h:382-446), with the L-2 norm of g
  // when l2 != nullptr.
  auto run = [&](auto &sp, double *x, auto &&eval, auto &&stat) {
    auto check = [&](int step, const double *g, double &cv) -> int {
      double l2 = 0.0;
      cv = stat(g, 1.0, c->params.verbose_output ? &l2 : nullptr);
      if (c->params.verbose_output) report(step, l2, cv);
      if (cv <= tol) return kSuccess;
      if ((unsigned)step >= max_it || std::isnan(cv)) return kFailure;
      return kIterate;
    };
    if (low) {
      // FP32_LSE: iterate on FP32 evaluations until the stopping value is within
      // kPolishFactor of the tolerance (or the budget ends, or a line search
      // fails), then finish on FP64 evaluations, so that the stopping decision
      // is taken on an FP64 gradient
      auto check_low = [&](int step, const double *g, double &cv) -> int {
        if (c->params.verbose_output) {
          double l2 = 0.0;
          const double v = stat(g, 1.0, &l2);
          report(step, l2, v);
        }
        cv = stat(g, kPolishFactor, nullptr);
        if (cv <= (comp ? DBL_MIN : utol * kPolishFactor)) return kSuccess;
        if ((unsigned)step >= max_it || std::isnan(cv)) return kFailure;
        return kIterate;
      };
      r = lbfgs_core(sp, x, eval, check_low, 5, noise_aware ? kNoiseRelF32 : 0.0);
      if (eval_rc) return;
      its = r.iterations;
      evals = r.evaluations;
      if (r.status != kNonFinite) {
        const unsigned left = max_it > (unsigned)its ? max_it - (unsigned)its : 0u;
        const int offset = its;
        auto check_hi = [&](int step, const double *g, double &cv) -> int {
          double l2 = 0.0;
          cv = stat(g, 1.0, c->params.verbose_output ? &l2 : nullptr);
          if (c->params.verbose_output) report(offset + step, l2, cv);
          if (cv <= tol) return kSuccess;
          if ((unsigned)step >= left || std::isnan(cv)) return kFailure;
          return kIterate;
        };
        c->force_mode = kF64Lse;
        c->hist_valid = false;
        r = lbfgs_core(sp, x, eval, check_hi, 5, noise_aware ? kNoiseRelF64 : 0.0);
        c->force_mode = -1;
        if (eval_rc) return;
        its += r.iterations;
        evals += r.evaluations;
      }
    } else {
      r = lbfgs_core(sp, x, eval, check, 5, noise_aware ? kNoiseRelF64 : 0.0);
      if (eval_rc) return;
      its = r.iterations;
      evals = r.evaluations;
    }
  };
  if (N >= device_pairs_min()) {
    // every vector in HBM; psi goes up once and comes back once
    DeviceSpace sp;
    auto space_rc = [&]() {
      return fail(SDOT_ERR_HIP, std::string("L-BFGS on the device: ") + sp.what + ": " +
                                    hipGetErrorString(sp.err));
    };
    if (!sp.init(N, 5, h0, c->h_nu.data(), c->stream) || !sp.upload(sp.x(), psi, "psi upload"))
      return space_rc();
    auto eval = [&](const double *x, double *g) -> double {
      double F = NAN;
      if (eval_rc == SDOT_OK) {
        eval_rc = stateful_eval_device(c, x, g, sp.part(), sp.hpart(), &F);
        if (eval_rc) F = NAN;
      }
      return F;
    };
    auto stat = [&](const double *g, double mult, double *l2) -> double {
      double st[kLbStats];
      if (eval_rc || !sp.stats(g, utol * mult, st)) return NAN;
      if (l2) *l2 = std::sqrt(st[kLbL2]);
      return comp ? st[kLbCw] : st[kLbL1];
    };
    run(sp, sp.x(), eval, stat);
    const bool back = sp.download(psi, sp.x(), "psi download");
    if (eval_rc) {
      c->force_mode = -1;
      return eval_rc;
    }
    if (!back || !sp.ok()) return space_rc();
  } else {
    HostSpace sp(N, 5, h0);
    auto eval = [&](const double *x, double *g) -> double {
      double F = NAN;
      if (eval_rc == SDOT_OK) {
        eval_rc = stateful_eval(c, x, g, &F, nullptr);
        if (eval_rc) {
          F = NAN;
          for (int64_t j = 0; j < N; ++j) g[j] = NAN;
        }
      }
      return F;
    };
    auto stat = [&](const double *g, double mult, double *l2) -> double {
      if (l2) {
        double q = 0.0;
        for (int64_t j = 0; j < N; ++j) q += g[j] * g[j];
        *l2 = std::sqrt(q);
      }
      double cv;
      if (comp) {
        cv = -INFINITY;
        for (int64_t j = 0; j < N; ++j) cv = std::max(cv, std::fabs(g[j]) - c->h_nu[j] * utol * mult);
      } else {
        cv = 0.0;
        for (int64_t j = 0; j < N; ++j) cv += std::fabs(g[j]);
      }
      return cv;
    };
    run(sp, psi, eval, stat);
    if (eval_rc) {
      c->force_mode = -1;
      return eval_rc;
    }
  }
  c->last_iterations = its;
  c->converged = r.status == kConverged;
  if (info) {
    info->status = r.status;
    info->iterations = its;
    info->evaluations = evals;
    info->functional = r.f;
    info->check_value = r.check;
    info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return SDOT_OK;
}

extern "C" int sdot_covering_radius(sdot_ctx *c, double *r0) {
  if (ctx_check(c) || !r0) return fail(SDOT_ERR_INVALID, "null argument");
  if (!c->have_source || !c->have_target) return fail(SDOT_ERR_STATE, "measures not set");
  HIP_TRY(hipSetDevice(c->device));
  return raw_covering_radius(c, r0);
}

extern "C" int sdot_distance_threshold(sdot_ctx *c, const double *psi, double *R) {
  if (ctx_check(c) || !psi || !R) return fail(SDOT_ERR_INVALID, "null argument");
  if (!c->have_source || !c->have_target) return fail(SDOT_ERR_STATE, "measures not set");
  HIP_TRY(hipSetDevice(c->device));
  return threshold(c, psi, R);
}

extern "C" int sdot_set_distance_threshold(sdot_ctx *c, double R) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  c->current_threshold = R;
  return SDOT_OK;
}

extern "C" int sdot_candidate_stats(sdot_ctx *c, double R, int64_t *count, uint64_t *hash,
                                    int64_t *pairs) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  HIP_TRY(hipSetDevice(c->device));
  int rc = prepare(c, R);
  if (rc) return rc;
  const Plan &p = c->plan;
  // target records carry the raw coordinates the candidate rule reads
  std::vector<double> zero(p.N, 0.0);
  HIP_TRY(hipMemcpy(c->dev.psi, zero.data(), sizeof(double) * p.N, hipMemcpyHostToDevice));
  rc = run_eval(c, R, c->dev.psi, c->dev.grad, c->dev.out);
  if (rc) return rc;
  int64_t *d_cnt = nullptr;
  uint64_t *d_hash = nullptr;
  HIP_TRY(hipMalloc((void **)&d_cnt, sizeof(int64_t) * p.n_cells));
  HIP_TRY(hipMalloc((void **)&d_hash, sizeof(uint64_t) * p.n_cells));
  hipError_t e = launch_candidate_stats(p, c->dev, d_cnt, d_hash, c->stream);
  std::vector<int64_t> hc(p.n_cells);
  std::vector<uint64_t> hh(p.n_cells);
  if (e == hipSuccess)
    e = hipMemcpyAsync(hc.data(), d_cnt, sizeof(int64_t) * p.n_cells, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(hh.data(), d_hash, sizeof(uint64_t) * p.n_cells, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d_cnt);
  (void)hipFree(d_hash);
  if (e != hipSuccess) return fail(SDOT_ERR_HIP, std::string("candidate stats: ") + hipGetErrorString(e));
  int64_t tot = 0;
  for (int64_t i = 0; i < p.n_cells; ++i) tot += hc[i];
  if (count) std::memcpy(count, hc.data(), sizeof(int64_t) * p.n_cells);
  if (hash) std::memcpy(hash, hh.data(), sizeof(uint64_t) * p.n_cells);
  if (pairs) *pairs = tot * p.nq;
  c->last_pairs = tot * p.nq;
  return SDOT_OK;
}

extern "C" int sdot_q_log_partition(sdot_ctx *c, double R, const double *psi, double *logZ) {
  if (ctx_check(c) || !psi || !logZ) return fail(SDOT_ERR_INVALID, "null argument");
  if (!c->have_source || !c->have_target) return fail(SDOT_ERR_STATE, "measures not set");
  HIP_TRY(hipSetDevice(c->device));
  c->force_mode = kF64Lse;
  int rc = host_eval(c, R, psi, nullptr, nullptr, nullptr);
  c->force_mode = -1;
  if (rc) return rc;
  const Plan &p = c->plan;
  std::vector<double> rec((size_t)p.nQ * kRec);
  HIP_TRY(hipMemcpy(rec.data(), c->dev.qrec_r, sizeof(double) * rec.size(), hipMemcpyDeviceToHost));
  for (int64_t cr = 0; cr < p.n_cells; ++cr) {
    const int64_t co = c->h_cell_r2o[cr];
    for (int k = 0; k < p.nq; ++k) {
      const double *r = &rec[(size_t)(cr * p.nq + k) * kRec];
      const double M = r[p.dim], scale = r[p.dim + 1], S = r[p.dim + 2];
      logZ[co * p.nq + k] = scale > 0.0 ? M + std::log(S) : -INFINITY;
    }
  }
  return SDOT_OK;
}

extern "C" int sdot_get_state(sdot_ctx *c, sdot_state *st) {
  if (ctx_check(c) || !st) return fail(SDOT_ERR_INVALID, "null argument");
  st->last_functional = c->global_functional;
  st->C_global = c->C_global;
  st->last_distance_threshold = c->current_threshold;
  st->covering_radius = c->covering_radius;
  st->min_target_density = c->min_target_density;
  st->last_iteration_count = c->last_iterations;
  st->converged = c->converged;
  st->last_pair_count = c->last_pairs;
  return SDOT_OK;
}

extern "C" int sdot_set_state(sdot_ctx *c, const sdot_state *st) {
  if (ctx_check(c) || !st) return fail(SDOT_ERR_INVALID, "null argument");
  c->global_functional = st->last_functional;
  c->C_global = st->C_global;
  c->current_threshold = st->last_distance_threshold;
  c->covering_radius = st->covering_radius;
  c->min_target_density = st->min_target_density;
  return SDOT_OK;
}

extern "C" int sdot_profile_enable(sdot_ctx *c, int on) {
  if (ctx_check(c)) return SDOT_ERR_INVALID;
  HIP_TRY(hipSetDevice(c->device));
  if (on && c->ev_pool.empty()) {
    c->ev_pool.resize(kProfSets);
    for (auto &set : c->ev_pool)
      for (auto &e : set) HIP_TRY(hipEventCreate(&e));
  }
  int rc = resolve_profile(c);
  if (rc) return rc;
  c->prof_on = on != 0;
  c->prof_acc = sdot_profile{};
  return SDOT_OK;
}

extern "C" int sdot_profile_get(sdot_ctx *c, sdot_profile *out) {
  if (ctx_check(c) || !out) return fail(SDOT_ERR_INVALID, "null argument");
  int rc = resolve_profile(c);
  if (rc) return rc;
  *out = c->prof_acc;
  return SDOT_OK;
}

// ---------------------------------------------------------------------------
// source arrays lent to the serving queue (sdot_serve.cpp)
// ---------------------------------------------------------------------------
int sdot::ctx_source_view(sdot_ctx *c, SourceView &v) {
  if (!c) return fail(SDOT_ERR_INVALID, "null context");
  if (!c->have_source) return fail(SDOT_ERR_STATE, "the context has no source measure");
  const Plan &p = c->plan;
  const Dev &d = c->dev;
  v.device = c->device;
  v.dim = p.dim;
  v.nq = p.nq;
  v.G = p.grp_big.G;
  v.gpc = p.grp_big.gpc;
  v.n_cells = p.n_cells;
  v.nQ = p.nQ;
  v.n_groups = p.grp_big.n_groups;
  v.n_blocks = p.grp_big.n_blocks;
  v.cg = p.cg;
  for (int k = 0; k < 3; ++k) {
    v.origin[k] = p.origin[k];
    v.lo[k] = c->src_lo[k];
    v.hi[k] = c->src_hi[k];
  }
  v.coord_scale = c->coord_scale;
  v.qx_m = d.qx_m;
  v.qw_m = d.qw_m;
  v.qx2_m = d.qx2_m;
  v.cc_m = d.cc_m;
  v.blkbox = d.blkbox;
  v.cc_r = d.cc_r;
  v.cstart = d.cstart;
  v.cell_r2m = d.cell_r2m;
  v.Gs = p.grp_small.G;
  v.gpc_s = p.grp_small.gpc;
  v.n_groups_s = p.grp_small.n_groups;
  v.n_blocks_s = p.grp_small.n_blocks;
  v.n_cbins = p.n_cbins;
  v.blkbox_s = d.blkbox_s;
  v.qbox_s = d.qbox_s;
  v.rq = p.rq;
  return SDOT_OK;
}

void sdot::ctx_attach(sdot_ctx *c, int delta) { c->n_servers += delta; }
