// sdot_serve.cpp -- continuous-batching solve queue (include/sdot.h, "batched
// serving").  Host side of the serving path:
//   * admission: requests wait in submission order and are admitted while a
//     slot is free and their buffers fit the device arena (one hipMalloc at
//     creation, first-fit sub-allocation, so admission and retirement never
//     call hipMalloc / hipFree on the hot path);
//   * per solve: argument blocks of every evaluation pass (the same blocks the
//     single-solve launcher builds), the potential, gradient and L-BFGS
//     history in the arena, and the line-search state machine on the host
//     (LineSearchSM, the decisions of SotSolver::solve's deal.II line search,
//     SotSolver.cc:190-198);
//   * one step: upload the batch's actions, replay the captured step graph
//     (sdot_serve_kernels.hip), read back F, slopes and stopping statistics,
//     advance every solve's state machine; finished solves are retired at once
//     and their slots refilled at the next step.
// A solve's kernel arguments and segment counts depend only on its own
// request, so its arithmetic is the same alone and in any batch.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "sdot_bfgs.hpp"
#include "sdot_serve.hpp"

using namespace sdot;

#define SV_TRY(expr)                                                                      \
  do {                                                                                    \
    hipError_t _e = (expr);                                                               \
    if (_e != hipSuccess)                                                                 \
      return set_error(SDOT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));  \
  } while (0)

namespace {

constexpr size_t kAlign = 256;
constexpr int kMaxSlots = 4096;
constexpr int kMaxSegServe = 128;
constexpr int kDefaultPassTargets = 2048;
// 64 target-major items per solve on the 262,144-point mesh: the first 8
// requests of the serving trace ran 5.82 s against 6.06 s with 16384
constexpr int kDefaultGradQ = 4096;

size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over one device allocation (offsets, 256-byte aligned).
class Arena {
 public:
  hipError_t init(size_t bytes) {
    size_ = align_up(bytes);
    hipError_t e = hipMalloc(&base_, size_);
    if (e == hipSuccess) free_[0] = size_;
    return e;
  }
  void release() {
    if (base_) (void)hipFree(base_);
    base_ = nullptr;
    free_.clear();
  }
  size_t alloc(size_t bytes) {
    bytes = align_up(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, rem = it->second - bytes;
      free_.erase(it);
      if (rem) free_[off + bytes] = rem;
      used_ += bytes;
      return off;
    }
    return SIZE_MAX;
  }
  void release_block(size_t off, size_t bytes) {
    bytes = align_up(bytes);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  char *at(size_t off) const { return static_cast<char *>(base_) + off; }
  size_t size() const { return size_; }
  size_t used() const { return used_; }

 private:
  void *base_ = nullptr;
  size_t size_ = 0, used_ = 0;
  std::map<size_t, size_t> free_;
};

// Byte offsets of one solve's buffers inside its arena block.
struct Carve {
  size_t off = 0;
  template <class T>
  size_t take(size_t n) {
    const size_t o = off;
    off = align_up(off + std::max<size_t>(n, 1) * sizeof(T));
    return o;
  }
};

struct SolveShape {
  int64_t N = 0, ntb = 0, ntr = 0, n_twaves = 0, nPpart = 0;
  size_t lay_bytes = 0;  // scratch of the device-built target layout
  int K1 = 1, K2 = 1;
  bool fp32 = false;
};

struct SolveOffsets {
  size_t y, nu, pos_r, pos_m, tm2o, tstart, twbox;
  size_t trec_r, trec_m, qrec_r, qchunk, Mpart, Spart, Tpart, Gpart;
  size_t Fpart, Cpart, Upart, PNpart, psi_prev, dpsi;
  size_t psit_r, pbin, prow, vcell, vbin, vrow;  // value bound (FP32 solves)
  size_t x, g, x0, gv, p, S[kHist], Y[kHist], hist;
  size_t total;
};

SolveOffsets carve(const SourceView &src, const SolveShape &s) {
  Carve c;
  SolveOffsets o{};
  const int64_t nFpart = (src.nQ + 255) / 256;
  o.y = c.take<double>((size_t)s.N * src.dim);
  o.nu = c.take<double>(s.N);
  o.pos_r = c.take<int32_t>(s.N);
  o.pos_m = c.take<int32_t>(s.N);
  o.tm2o = c.take<int32_t>(s.N);
  o.tstart = c.take<int32_t>(s.ntb + 1);
  o.twbox = c.take<double>((size_t)s.n_twaves * 2 * src.dim);
  o.trec_r = c.take<double>((size_t)s.N * kRec);
  o.trec_m = c.take<double>((size_t)s.N * kRec);
  o.qrec_r = c.take<double>((size_t)src.nQ * kRec);
  o.qchunk = c.take<double>((size_t)src.nQ * (src.dim + 2));
  o.Mpart = c.take<double>((size_t)s.K1 * src.nQ);
  o.Spart = c.take<double>((size_t)s.K1 * src.nQ);
  o.Tpart = s.fp32 ? c.take<double>((size_t)s.K1 * src.nQ) : SIZE_MAX;
  o.Gpart = c.take<double>((size_t)s.K2 * s.N);
  o.Fpart = c.take<double>(nFpart);
  o.Cpart = c.take<double>(nFpart);
  o.Upart = c.take<double>(nFpart);
  o.PNpart = c.take<double>(s.nPpart);
  o.psi_prev = s.fp32 ? c.take<double>(s.N) : SIZE_MAX;
  o.dpsi = s.fp32 ? c.take<double>(s.N) : SIZE_MAX;
  const int64_t ncr = grid_rows(src.cg);
  o.psit_r = s.fp32 ? c.take<double>(s.N) : SIZE_MAX;
  o.pbin = s.fp32 ? c.take<double>(s.ntb) : SIZE_MAX;
  o.prow = s.fp32 ? c.take<double>(s.ntr) : SIZE_MAX;
  o.vcell = s.fp32 ? c.take<double>(src.n_cells) : SIZE_MAX;
  o.vbin = s.fp32 ? c.take<double>(src.n_cbins) : SIZE_MAX;
  o.vrow = s.fp32 ? c.take<double>(ncr) : SIZE_MAX;
  o.x = c.take<double>(s.N);
  o.g = c.take<double>(s.N);
  o.x0 = c.take<double>(s.N);
  o.gv = c.take<double>(s.N);
  o.p = c.take<double>(s.N);
  for (int k = 0; k < kHist; ++k) o.S[k] = c.take<double>(s.N);
  for (int k = 0; k < kHist; ++k) o.Y[k] = c.take<double>(s.N);
  o.hist = c.take<SlotHist>(1);
  // the target layout is built at admission in the space of the evaluation
  // buffers (trec_r onwards), which the first evaluation writes afterwards
  o.total = std::max(c.off, align_up(o.trec_r + s.lay_bytes));
  return o;
}

struct Request {
  int64_t id = 0;
  int64_t N = 0;
  double eps = 1.0, tol = 0.0, R = INFINITY;
  std::vector<double> eps_seq;  // epsilon schedule, largest first (last = eps)
  int control = SDOT_CTRL_L1NORM;
  int line_search = SDOT_LS_REFERENCE;
  int precond = SDOT_PRECOND_NONE;
  uint32_t max_it = 0;
  std::vector<double> y, nu, psi0;
  TargetShape L;
  SolveShape shape;
  int dense = 0;
  double smax = 0.0, Rq = 0.0, s_in = 0.0;
  double diag = 0.0, nu_min = 0.0;  // source-box diagonal, smallest weight (value bound)
  size_t bytes = 0;
  double t_submit = 0.0;
};

struct Solve {
  std::unique_ptr<Request> req;
  size_t off = SIZE_MAX;
  SolveOffsets o{};
  SlotDesc desc{};
  SlotVecs vecs{};
  int64_t items[kNumLists] = {};
  int mode = kF32Lse;   // arithmetic of this solve's evaluations
  SolveDriver drv;
  size_t stage = 0;     // index into the epsilon schedule
  int it_done = 0, ev_done = 0;  // iterations / evaluations of finished stages
  bool seeded = false;  // FP32: the next evaluation starts from the history seed
  double t_admit = 0.0;
};

struct Done {
  sdot_result res{};
  std::vector<double> psi;
};

}  // namespace

struct sdot_server {
  sdot_ctx *ctx = nullptr;
  SourceView src;
  int device = 0;
  int mode = kF32Lse;   // mode a solve starts in
  double polish = 10.0;
  int max_slots = 256;
  int use_graph = 1;
  int pass_targets = kDefaultPassTargets, grad_qpoints = kDefaultGradQ;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float last_step_ms = 0.0f;
  Arena arena;
  // batch block: device copy and pinned host staging, same layout
  //   [desc][vecs][base 5 x bstride][in][nact]
  size_t off_desc = 0, off_vecs = 0, off_base = 0, off_in = 0, off_nact = 0, block_bytes = 0;
  char *d_block = nullptr, *h_block = nullptr;
  SlotOut *d_out = nullptr, *h_out = nullptr;
  int *d_ctr = nullptr;
  int64_t bstride = 0;
  ServeDev sd{};
  std::vector<std::unique_ptr<Solve>> active;
  bool dirty = true;
  // shared with submit / poll
  std::mutex mu;
  std::deque<std::unique_ptr<Request>> queue;
  std::unordered_map<int64_t, Done> done;
  int64_t next_id = 0;
  std::chrono::steady_clock::time_point t0;

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

int fill_shape(sdot_server *s, Request &r) {
  const SourceView &src = s->src;
  SolveShape &sh = r.shape;
  sh.N = r.N;
  sh.ntb = r.L.n_bins;
  sh.ntr = grid_rows(r.L.tg);
  sh.n_twaves = r.L.n_twaves;
  sh.lay_bytes = layout_scratch_bytes(r.L);
  sh.nPpart = (r.N + 255) / 256;
  sh.fp32 = s->mode == kF32Lse;
  // candidate rule and enumeration radius, as for a single solve
  double diag2 = 0.0;
  for (int d = 0; d < src.dim; ++d) {
    const double lo = std::min(src.lo[d], r.L.lo[d]);
    const double hi = std::max(src.hi[d], r.L.hi[d]);
    diag2 += (hi - lo) * (hi - lo);
  }
  const double diag = std::sqrt(diag2);
  double sdiag2 = 0.0;
  for (int d = 0; d < src.dim; ++d) sdiag2 += (src.hi[d] - src.lo[d]) * (src.hi[d] - src.lo[d]);
  r.diag = std::sqrt(sdiag2);  // the value bound's default rule takes the source's extent
  r.nu_min = r.nu.empty() ? 0.0 : *std::min_element(r.nu.begin(), r.nu.end());
  const double R = r.R;
  r.smax = smax_for(R);
  r.s_in = inner_s(r.smax);
  r.dense = (R >= diag * (1.0 + 1e-6) + 1e-300) ? 1 : 0;
  r.Rq = R * (1.0 + 1e-9) + 1e-9 * (r.L.tg.h + src.cg.h + std::max(src.coord_scale, r.L.scale));
  if (!std::isfinite(r.Rq)) r.dense = 1;
  // segment counts from the solve's own size: work items of about
  // pass_targets targets (q-major) and grad_qpoints q-points (target-major)
  const int64_t k1 = (r.N + s->pass_targets - 1) / s->pass_targets;
  const int64_t k2 = (src.nQ + s->grad_qpoints - 1) / s->grad_qpoints;
  sh.K1 = (int)std::clamp<int64_t>(k1, 1, kMaxSegServe);
  sh.K2 = (int)std::clamp<int64_t>(k2, 1, kMaxSegServe);
  // work items run segment-fastest and every solve's item count is then a
  // multiple of 8, so a persistent workgroup (fixed XCD) sees the segments
  // s = xcd (mod 8) only and its L2 keeps one eighth of the q records
  sh.K2 = (sh.K2 + 7) / 8 * 8;
  r.bytes = carve(src, sh).total;
  return SDOT_OK;
}

template <class T>
T *dptr(sdot_server *s, const Solve &v, size_t rel) {
  return rel == SIZE_MAX ? nullptr : reinterpret_cast<T *>(s->arena.at(v.off + rel));
}

// Argument blocks of the solve's evaluation (as the single-solve launcher
// builds them), pointing into its arena block.
void build_desc(sdot_server *s, Solve &v) {
  const SourceView &src = s->src;
  const Request &r = *v.req;
  const SolveShape &sh = r.shape;
  const double eps = r.eps_seq[v.stage];
  const bool fp32 = v.mode == kF32Lse;
  const double coef = (1.0 / eps) * (fp32 ? 1.44269504088896340736 : 1.0);
  const double ln_scale = fp32 ? 0.69314718055994530942 : 1.0;
  double *trec_r = dptr<double>(s, v, v.o.trec_r), *trec_m = dptr<double>(s, v, v.o.trec_m);
  double *qrec_r = dptr<double>(s, v, v.o.qrec_r);
  double *Mpart = dptr<double>(s, v, v.o.Mpart), *Spart = dptr<double>(s, v, v.o.Spart);
  double *Gpart = dptr<double>(s, v, v.o.Gpart);
  double *Fpart = dptr<double>(s, v, v.o.Fpart), *Cpart = dptr<double>(s, v, v.o.Cpart);
  double *Upart = dptr<double>(s, v, v.o.Upart), *PNpart = dptr<double>(s, v, v.o.PNpart);
  double *dpsi = dptr<double>(s, v, v.o.dpsi);
  double *x = dptr<double>(s, v, v.o.x), *g = dptr<double>(s, v, v.o.g);
  const int64_t nFpart = (src.nQ + 255) / 256;
  // value bound: the single-solve default rule at this stage's epsilon
  const bool bound = fp32 && r.nu_min > 0.0 && value_bound_default(eps, r.diag, r.R, r.N, src.nQ);
  const double theta = kPruneBitsF32;  // FP32 sweeps: log2 units
  double *psit_r = dptr<double>(s, v, v.o.psit_r);
  double *pbin = dptr<double>(s, v, v.o.pbin), *prow = dptr<double>(s, v, v.o.prow);
  double *vcell = dptr<double>(s, v, v.o.vcell), *vbin = dptr<double>(s, v, v.o.vbin);
  double *vrow = dptr<double>(s, v, v.o.vrow);

  PrepArgs &pa = v.desc.prep;
  pa = PrepArgs{};
  pa.y = dptr<double>(s, v, v.o.y);
  pa.nu = dptr<double>(s, v, v.o.nu);
  pa.psi = x;
  pa.pos_r = dptr<int32_t>(s, v, v.o.pos_r);
  pa.pos_m = dptr<int32_t>(s, v, v.o.pos_m);
  pa.trec_r = trec_r;
  pa.trec_m = trec_m;
  pa.PNpart = PNpart;
  pa.psi_prev = fp32 ? dptr<double>(s, v, v.o.psi_prev) : nullptr;
  pa.dpsi_r = dpsi;
  pa.N = r.N;
  pa.hist = v.seeded ? 1 : 0;
  for (int k = 0; k < 3; ++k) pa.o[k] = src.origin[k];
  pa.coef = coef;
  pa.eps = eps;
  pa.psit_r = bound ? psit_r : nullptr;

  P1Args &a = v.desc.p1;
  a = P1Args{};
  a.qx = src.qx_m;
  a.qx2 = src.qx2_m;
  a.cc = src.cc_m;
  a.blkbox = src.blkbox;
  a.trec = trec_r;
  a.tstart = dptr<int32_t>(s, v, v.o.tstart);
  a.Mpart = Mpart;
  a.dpsi = dpsi;
  a.Mout = Mpart;
  a.Sout = Spart;
  a.Tout = dptr<double>(s, v, v.o.Tpart);
  a.tg = r.L.tg;
  a.n_groups = src.n_groups;
  a.nQ = src.nQ;
  a.N = r.N;
  a.nq = src.nq;
  a.gpc = src.gpc;
  a.K1 = sh.K1;
  a.dense = r.dense;
  a.hist = v.seeded ? 1 : 0;
  a.hist_same = 1;  // a solve's radius does not change
  a.smax = r.smax;
  a.Rq = r.Rq;
  a.s_in = r.s_in;
  a.plain_coef = 0.5 / eps;
  a.coef = coef;
  if (bound) {
    // the 4-point grouping (blocks of half as many cells) and the bound
    a.blkbox = src.blkbox_s;
    a.n_groups = src.n_groups_s;
    a.gpc = src.gpc_s;
    a.prune = 1;
    a.qbox = src.qbox_s;
    a.pbin = pbin;
    a.prow = prow;
    for (int k = 0; k < 3; ++k) a.o[k] = src.origin[k];
    a.theta = theta;
    a.inv_coef = 1.0 / coef;
    a.vshift = eps * std::log(r.nu_min);
  }

  QFArgs &q = v.desc.qf;
  q = QFArgs{};
  q.qx = src.qx_m;
  q.qw = src.qw_m;
  q.qx2 = src.qx2_m;
  q.Mpart = Mpart;
  q.Spart = Spart;
  q.Tpart = dptr<double>(s, v, v.o.Tpart);
  q.cell_r2m = src.cell_r2m;
  q.qrec_r = qrec_r;
  q.qchunk = dptr<double>(s, v, v.o.qchunk);
  q.G = (fp32 && bound) ? src.Gs : src.G;  // the chunk size of this solve's kl_grad
  q.Fpart = Fpart;
  q.Cpart = Cpart;
  q.Upart = Upart;
  q.nQ = src.nQ;
  q.nq = src.nq;
  q.K1 = sh.K1;
  q.eps = eps;
  q.plain_coef = 0.5 / eps;
  q.ln_scale = ln_scale;

  GArgs &gr = v.desc.gr;
  gr = GArgs{};
  gr.trec_m = trec_m;
  gr.twbox = dptr<double>(s, v, v.o.twbox);
  gr.cc_r = src.cc_r;
  gr.cstart = src.cstart;
  gr.qrec_r = qrec_r;
  gr.qchunk = q.qchunk;
  gr.Gpart = Gpart;
  gr.cg = src.cg;
  gr.N = r.N;
  gr.n_cells = src.n_cells;
  gr.nq = src.nq;
  gr.K2 = sh.K2;
  gr.dense = r.dense;
  gr.smax = r.smax;
  gr.Rq = r.Rq;
  gr.s_in = r.s_in;
  if (bound) {
    gr.prune = 1;
    gr.vcell = vcell;
    gr.vbin = vbin;
    gr.vrow = vrow;
    gr.rq = src.rq;
    gr.inv_coef = 1.0 / coef;
    gr.eps = eps;
  }

  VBArgs &vb = v.desc.vb;
  vb = VBArgs{};
  vb.psit_r = psit_r;
  vb.tstart = a.tstart;
  vb.pbin = pbin;
  vb.prow = prow;
  vb.ntb = sh.ntb;
  vb.ntr = sh.ntr;
  vb.tnx = r.L.tg.n[0];
  vb.qrec_r = qrec_r;
  vb.cstart = src.cstart;
  vb.vcell = vcell;
  vb.vbin = vbin;
  vb.vrow = vrow;
  vb.n_cells = src.n_cells;
  vb.ncb = src.n_cbins;
  vb.ncr = grid_rows(src.cg);
  vb.cnx = src.cg.n[0];
  vb.nq = src.nq;
  vb.log2_units = 1;
  vb.theta = theta;
  vb.inv_coef = 1.0 / coef;

  v.desc.gf = GFArgs{Gpart, trec_m, dptr<int32_t>(s, v, v.o.tm2o), g, r.N, sh.K2};
  // out / ucount are set per batch position in upload_batch
  v.desc.fin = FinArgs{Fpart, Cpart, Upart, PNpart, nFpart, sh.nPpart, nullptr, nullptr};
  v.desc.nPpart = sh.nPpart;

  SlotVecs &w = v.vecs;
  w.x = x;
  w.g = g;
  w.x0 = dptr<double>(s, v, v.o.x0);
  w.gv = dptr<double>(s, v, v.o.gv);
  w.p = dptr<double>(s, v, v.o.p);
  for (int k = 0; k < kHist; ++k) {
    w.S[k] = dptr<double>(s, v, v.o.S[k]);
    w.Y[k] = dptr<double>(s, v, v.o.Y[k]);
  }
  w.nu = pa.nu;
  w.hist = dptr<SlotHist>(s, v, v.o.hist);
  w.N = r.N;

  for (int64_t &it : v.items) it = 0;
  v.items[kLPrep] = sh.nPpart;
  v.items[fp32 ? (bound ? kLPass32V : kLPass32) : kLPass64] =
      (bound ? src.n_blocks_s : src.n_blocks) * sh.K1;
  v.items[fp32 ? kLQfin32 : kLQfin64] = nFpart;
  v.items[fp32 ? (bound ? kLGrad32V : kLGrad32) : kLGrad64] = sh.n_twaves * sh.K2;
  v.items[kLGfin] = (r.N + 255) / 256;
  if (bound) {
    v.items[kLTbin] = (sh.ntb + 3) / 4;
    v.items[kLTrow] = (sh.ntr + 255) / 256;
    v.items[kLCell] = (src.n_cells + 255) / 256;
    v.items[kLCbin] = (src.n_cbins + 3) / 4;
    v.items[kLCrow] = (vb.ncr + 255) / 256;
  }
}

const SlotHist kEmptyHist = [] {
  SlotHist h{};
  h.count = 0;
  h.newest = -1;
  return h;
}();

void init_driver(sdot_server *s, Solve &v) {
  const Request &r = *v.req;
  v.mode = s->mode;
  v.seeded = false;
  const bool aware = r.line_search == SDOT_LS_NOISE_AWARE;
  const double noise_rel = aware ? (v.mode == kF32Lse ? kNoiseRelF32 : kNoiseRelF64) : 0.0;
  v.drv = SolveDriver();
  v.drv.init(r.control, r.tol, r.max_it, noise_rel, v.mode == kF32Lse ? s->polish : 0.0,
             aware ? kNoiseRelF64 : 0.0);
}

// Uploads an admitted request into its arena block (offset `off`).
int admit(sdot_server *s, std::unique_ptr<Request> r, const SolveOffsets &o, size_t off) {
  auto v = std::make_unique<Solve>();
  v->o = o;
  v->off = off;
  v->req = std::move(r);
  init_driver(s, *v);
  v->t_admit = s->now();
  Request &q = *v->req;
  const int dim = s->src.dim;
  auto up = [&](size_t rel, const void *h, size_t bytes) {
    return hipMemcpyAsync(s->arena.at(v->off + rel), h, bytes, hipMemcpyHostToDevice, s->stream);
  };
  SV_TRY(up(v->o.y, q.y.data(), sizeof(double) * q.N * dim));
  SV_TRY(up(v->o.nu, q.nu.data(), sizeof(double) * q.N));
  {
    auto at = [&](size_t rel) { return s->arena.at(v->off + rel); };
    SV_TRY(device_target_layout(dim, q.L, (const double *)at(v->o.y), (int32_t *)at(v->o.pos_r),
                                nullptr, (int32_t *)at(v->o.pos_m), (int32_t *)at(v->o.tm2o),
                                (int32_t *)at(v->o.tstart), (double *)at(v->o.twbox), nullptr,
                                at(v->o.trec_r), v->o.total - v->o.trec_r, s->stream));
  }
  SV_TRY(up(v->o.x, q.psi0.data(), sizeof(double) * q.N));
  SV_TRY(up(v->o.hist, &kEmptyHist, sizeof(SlotHist)));
  build_desc(s, *v);
  s->active.push_back(std::move(v));
  s->dirty = true;
  return SDOT_OK;
}

// Writes the batch block (staging) and uploads it.
int upload_batch(sdot_server *s) {
  const int n = (int)s->active.size();
  char *h = s->h_block;
  if (s->dirty) {
    auto *desc = reinterpret_cast<SlotDesc *>(h + s->off_desc);
    auto *vecs = reinterpret_cast<SlotVecs *>(h + s->off_vecs);
    auto *base = reinterpret_cast<int64_t *>(h + s->off_base);
    for (int k = 0; k < kNumLists; ++k) base[k * s->bstride] = 0;
    for (int i = 0; i < n; ++i) {
      Solve &v = *s->active[i];
      v.desc.prep.hist = v.seeded ? 1 : 0;
      v.desc.p1.hist = v.seeded ? 1 : 0;
      v.desc.fin.out = &s->d_out[i].F;
      v.desc.fin.ucount = &s->d_out[i].lost;
      desc[i] = v.desc;
      vecs[i] = v.vecs;
      for (int k = 0; k < kNumLists; ++k)
        base[k * s->bstride + i + 1] = base[k * s->bstride + i] + v.items[k];
    }
    SV_TRY(hipMemcpyAsync(s->d_block, h, s->off_in, hipMemcpyHostToDevice, s->stream));
    s->dirty = false;
  }
  auto *in = reinterpret_cast<SlotIn *>(h + s->off_in);
  for (int i = 0; i < n; ++i) {
    const Solve &v = *s->active[i];
    in[i] = v.drv.next_in();
    // diag(eps / nu_j) at the epsilon of the solve's current stage
    in[i].h0 = v.req->precond == SDOT_PRECOND_DIAGONAL ? v.req->eps_seq[v.stage] : 0.0;
  }
  *reinterpret_cast<int *>(h + s->off_nact) = n;
  SV_TRY(hipMemcpyAsync(s->d_block + s->off_in, h + s->off_in, s->block_bytes - s->off_in,
                        hipMemcpyHostToDevice, s->stream));
  return SDOT_OK;
}

int retire(sdot_server *s, size_t i, int status) {
  Solve &v = *s->active[i];
  const Request &r = *v.req;
  Done d;
  d.psi.resize(r.N);
  SV_TRY(hipMemcpyAsync(d.psi.data(), v.vecs.x, sizeof(double) * r.N, hipMemcpyDeviceToHost,
                        s->stream));
  SV_TRY(hipStreamSynchronize(s->stream));
  d.res.id = r.id;
  d.res.status = status;
  d.res.iterations = v.it_done + v.drv.iterations();
  d.res.evaluations = v.ev_done + v.drv.evaluations();
  d.res.functional = v.drv.functional();
  d.res.check_value = v.drv.check_value();
  d.res.rel_marginal_error = v.drv.relerr();
  d.res.t_submit = r.t_submit;
  d.res.t_admit = v.t_admit;
  d.res.t_done = s->now();
  s->arena.release_block(v.off, v.o.total);
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->done.emplace(r.id, std::move(d));
  }
  s->active.erase(s->active.begin() + i);
  s->dirty = true;
  return SDOT_OK;
}

// One solve's reaction to the evaluation of this step: a status >= 0 when
// the solve has finished, -1 otherwise.
int advance(sdot_server *s, Solve &v, const SlotOut &o) {
  const bool was_seeded = v.seeded;
  if (v.mode == kF32Lse && !v.seeded) {
    v.seeded = true;  // this evaluation left psi and M_q behind
    s->dirty = true;
  }
  // a seeded sweep that lost range is repeated without the seed
  const bool repeat = was_seeded && o.lost > 0.0;
  if (repeat) {
    v.seeded = false;
    s->dirty = true;
  }
  const int st = v.drv.on_result(o, repeat);
  if (st == SolveDriver::kSwitch) {
    // finish in FP64: same buffers, FP64 argument blocks and work lists
    v.mode = kF64Lse;
    v.seeded = false;
    build_desc(s, v);
    s->dirty = true;
    return -1;
  }
  return st;
}

int launch_step(sdot_server *s) {
  if (s->exec) {
    SV_TRY(hipGraphLaunch(s->exec, s->stream));
  } else {
    SV_TRY(launch_serve_step(s->src.dim, s->src.G, s->src.Gs, s->sd, s->stream));
  }
  return SDOT_OK;
}

}  // namespace

namespace {
int server_init(sdot_server *s, const sdot_server_config &cfg);
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int sdot_server_config_default(sdot_server_config *cfg) {
  if (!cfg) return set_error(SDOT_ERR_INVALID, "cfg is null");
  cfg->precision = SDOT_PREC_FP32_LSE;
  cfg->max_slots = 256;
  cfg->hbm_budget_bytes = (int64_t)16 << 30;
  cfg->use_graph = 1;
  cfg->pass_targets = 0;
  cfg->grad_qpoints = 0;
  cfg->polish_factor = kPolishFactor;
  return SDOT_OK;
}

extern "C" int sdot_server_create(sdot_ctx *ctx, const sdot_server_config *cfg_in,
                                  sdot_server **out) {
  if (!out) return set_error(SDOT_ERR_INVALID, "out is null");
  *out = nullptr;
  sdot_server_config cfg;
  sdot_server_config_default(&cfg);
  if (cfg_in) cfg = *cfg_in;
  if (cfg.precision != SDOT_PREC_FP64 && cfg.precision != SDOT_PREC_FP32_LSE)
    return set_error(SDOT_ERR_INVALID, "unknown precision");
  if (cfg.max_slots < 1 || cfg.max_slots > kMaxSlots)
    return set_error(SDOT_ERR_INVALID, "max_slots must be in [1, 4096]");
  if (cfg.hbm_budget_bytes < ((int64_t)1 << 20))
    return set_error(SDOT_ERR_INVALID, "hbm_budget_bytes must be at least 1 MiB");
  if (cfg.pass_targets < 0 || cfg.grad_qpoints < 0)
    return set_error(SDOT_ERR_INVALID, "work-item sizes must be >= 0");
  if (!(cfg.polish_factor >= 0.0))
    return set_error(SDOT_ERR_INVALID, "polish_factor must be >= 0");
  SourceView src;
  int rc = ctx_source_view(ctx, src);
  if (rc) return rc;
  sdot_server *s = new sdot_server();
  s->src = src;
  rc = server_init(s, cfg);
  if (rc) {
    sdot_server_destroy(s);
    return rc;
  }
  s->ctx = ctx;
  ctx_attach(ctx, +1);
  *out = s;
  return SDOT_OK;
}

namespace {

int server_init(sdot_server *s, const sdot_server_config &cfg) {
  s->device = s->src.device;
  s->mode = cfg.precision == SDOT_PREC_FP32_LSE ? kF32Lse : kF64Lse;
  s->polish = cfg.polish_factor;
  s->max_slots = cfg.max_slots;
  s->use_graph = cfg.use_graph ? 1 : 0;
  if (cfg.pass_targets) s->pass_targets = cfg.pass_targets;
  if (cfg.grad_qpoints) s->grad_qpoints = cfg.grad_qpoints;
  s->t0 = std::chrono::steady_clock::now();
  SV_TRY(hipSetDevice(s->device));
  SV_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  SV_TRY(hipEventCreate(&s->ev0));
  SV_TRY(hipEventCreate(&s->ev1));
  SV_TRY(s->arena.init((size_t)cfg.hbm_budget_bytes));
  // batch block
  s->bstride = s->max_slots + 1;
  Carve c;
  s->off_desc = c.take<SlotDesc>(s->max_slots);
  s->off_vecs = c.take<SlotVecs>(s->max_slots);
  s->off_base = c.take<int64_t>((size_t)kNumLists * s->bstride);
  s->off_in = c.take<SlotIn>(s->max_slots);
  s->off_nact = c.take<int>(1);
  s->block_bytes = c.off;
  SV_TRY(hipMalloc((void **)&s->d_block, s->block_bytes));
  SV_TRY(hipHostMalloc((void **)&s->h_block, s->block_bytes, hipHostMallocDefault));
  std::memset(s->h_block, 0, s->block_bytes);
  SV_TRY(hipMemcpy(s->d_block, s->h_block, s->block_bytes, hipMemcpyHostToDevice));
  SV_TRY(hipMalloc((void **)&s->d_out, sizeof(SlotOut) * s->max_slots));
  SV_TRY(hipMalloc((void **)&s->d_ctr, sizeof(int) * kNumQueues * 8));
  SV_TRY(hipMemset(s->d_ctr, 0, sizeof(int) * kNumQueues * 8));
  SV_TRY(hipHostMalloc((void **)&s->h_out, sizeof(SlotOut) * s->max_slots, hipHostMallocDefault));
  ServeDev &sd = s->sd;
  sd.desc = reinterpret_cast<SlotDesc *>(s->d_block + s->off_desc);
  sd.vecs = reinterpret_cast<SlotVecs *>(s->d_block + s->off_vecs);
  sd.base = reinterpret_cast<int64_t *>(s->d_block + s->off_base);
  sd.bstride = s->bstride;
  sd.in = reinterpret_cast<SlotIn *>(s->d_block + s->off_in);
  sd.nact = reinterpret_cast<int *>(s->d_block + s->off_nact);
  sd.out = s->d_out;
  sd.ctr = s->d_ctr;
  SV_TRY(serve_grid_sizes(s->src.dim, s->src.G, s->src.Gs, s->device, sd));
  sd.grid_slots = (unsigned)s->max_slots;
  if (s->use_graph) {
    SV_TRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    hipError_t le = launch_serve_step(s->src.dim, s->src.G, s->src.Gs, sd, s->stream);
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(s->stream, &g);
    s->graph = g;
    if (le != hipSuccess) return set_error(SDOT_ERR_HIP, std::string("step capture: ") + hipGetErrorString(le));
    if (ce != hipSuccess) return set_error(SDOT_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    SV_TRY(hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0));
  }
  return SDOT_OK;
}

}  // namespace

extern "C" int sdot_server_destroy(sdot_server *s) {
  if (!s) return SDOT_OK;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->exec) (void)hipGraphExecDestroy(s->exec);
  if (s->graph) (void)hipGraphDestroy(s->graph);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->d_block) (void)hipFree(s->d_block);
  if (s->h_block) (void)hipHostFree(s->h_block);
  if (s->d_out) (void)hipFree(s->d_out);
  if (s->d_ctr) (void)hipFree(s->d_ctr);
  if (s->h_out) (void)hipHostFree(s->h_out);
  s->arena.release();
  if (s->stream) (void)hipStreamDestroy(s->stream);
  if (s->ctx) ctx_attach(s->ctx, -1);
  delete s;
  return SDOT_OK;
}

extern "C" int sdot_server_submit(sdot_server *s, const sdot_request *q, int64_t *id) {
  if (!s || !q) return set_error(SDOT_ERR_INVALID, "null argument");
  if (q->n < 1 || q->n >= (1ll << 31) - 1) return set_error(SDOT_ERR_INVALID, "n must be >= 1");
  if (!q->y || !q->nu) return set_error(SDOT_ERR_INVALID, "null target array");
  if (!(q->epsilon > 0) || !std::isfinite(q->epsilon))
    return set_error(SDOT_ERR_INVALID, "epsilon must be > 0");
  if (q->solver_control_type != SDOT_CTRL_L1NORM && q->solver_control_type != SDOT_CTRL_COMPONENTWISE)
    return set_error(SDOT_ERR_INVALID, "unknown solver_control_type");
  const int dim = s->src.dim;
  auto r = std::make_unique<Request>();
  r->N = q->n;
  r->eps = q->epsilon;
  r->tol = q->tolerance;
  r->control = q->solver_control_type;
  r->max_it = q->max_iterations;
  if (q->line_search != SDOT_LS_REFERENCE && q->line_search != SDOT_LS_NOISE_AWARE)
    return set_error(SDOT_ERR_INVALID, "unknown line_search");
  r->line_search = q->line_search;
  if (q->preconditioner != SDOT_PRECOND_NONE && q->preconditioner != SDOT_PRECOND_DIAGONAL)
    return set_error(SDOT_ERR_INVALID, "unknown preconditioner");
  r->precond = q->preconditioner;
  r->R = (q->radius <= 0.0 || q->radius >= DBL_MAX || std::isnan(q->radius)) ? INFINITY : q->radius;
  // epsilon schedule of EpsilonScalingHandler::generate_epsilon_sequence
  // (EpsilonScalingHandler.cc:30-41; factor <= 1 -> 2, steps 0 -> 1 as its
  // constructor does, :18-27)
  {
    uint32_t steps = q->epsilon_scaling_steps == 0 ? 1 : q->epsilon_scaling_steps;
    const double factor = q->epsilon_scaling_factor > 1.0 ? q->epsilon_scaling_factor : 2.0;
    if (steps > 64) return set_error(SDOT_ERR_INVALID, "epsilon_scaling_steps must be <= 64");
    for (uint32_t i = 0; i < steps; ++i)
      r->eps_seq.push_back(q->epsilon * std::pow(factor, (double)(steps - 1 - i)));
  }
  r->y.assign(q->y, q->y + q->n * dim);
  r->nu.assign(q->nu, q->nu + q->n);
  for (double v : r->y)
    if (!std::isfinite(v)) return set_error(SDOT_ERR_INVALID, "non-finite target point");
  for (double v : r->nu)
    if (!(v > 0) || !std::isfinite(v)) return set_error(SDOT_ERR_INVALID, "target weights must be > 0");
  if (q->psi0) {
    r->psi0.assign(q->psi0, q->psi0 + q->n);
  } else {
    r->psi0.assign(q->n, 0.0);
  }
  target_shape(dim, r->N, r->y.data(), r->L);
  int rc = fill_shape(s, *r);
  if (rc) return rc;
  if (r->bytes > s->arena.size())
    return set_error(SDOT_ERR_NOMEM, "request needs " + std::to_string(r->bytes) +
                                         " device bytes, more than the server's HBM budget");
  std::lock_guard<std::mutex> g(s->mu);
  r->id = s->next_id++;
  r->t_submit = s->now();
  if (id) *id = r->id;
  s->queue.push_back(std::move(r));
  return SDOT_OK;
}

extern "C" int sdot_server_step(sdot_server *s, int32_t *n_active, int32_t *n_queued) {
  if (!s) return set_error(SDOT_ERR_INVALID, "null server");
  SV_TRY(hipSetDevice(s->device));
  // admission: submission order, while a slot is free and the arena has room
  for (;;) {
    std::unique_ptr<Request> r;
    {
      std::lock_guard<std::mutex> g(s->mu);
      if (s->queue.empty() || (int)s->active.size() >= s->max_slots) break;
      r = std::move(s->queue.front());
      s->queue.pop_front();
    }
    const SolveOffsets o = carve(s->src, r->shape);
    const size_t off = s->arena.alloc(o.total);
    if (off == SIZE_MAX) {
      std::lock_guard<std::mutex> g(s->mu);
      s->queue.push_front(std::move(r));
      break;
    }
    int rc = admit(s, std::move(r), o, off);
    if (rc) {
      // a HIP failure while copying the request in or building its layout:
      // the request is dropped and its block returned
      s->arena.release_block(off, o.total);
      return rc;
    }
  }
  const int n = (int)s->active.size();
  if (n > 0) {
    int rc = upload_batch(s);
    if (rc) return rc;
    SV_TRY(hipEventRecord(s->ev0, s->stream));
    rc = launch_step(s);
    if (rc) return rc;
    SV_TRY(hipEventRecord(s->ev1, s->stream));
    SV_TRY(hipMemcpyAsync(s->h_out, s->d_out, sizeof(SlotOut) * n, hipMemcpyDeviceToHost, s->stream));
    SV_TRY(hipStreamSynchronize(s->stream));
    SV_TRY(hipEventElapsedTime(&s->last_step_ms, s->ev0, s->ev1));
    std::vector<int> status(n);
    for (int i = 0; i < n; ++i) {
      Solve &v = *s->active[i];
      status[i] = advance(s, v, s->h_out[i]);
      if (status[i] >= 0 && v.stage + 1 < v.req->eps_seq.size()) {
        // epsilon scaling: the next stage starts from this potential with a
        // fresh L-BFGS history, whatever this stage's outcome
        // (SemiDiscreteOT.cc:888-917)
        v.it_done += v.drv.iterations();
        v.ev_done += v.drv.evaluations();
        ++v.stage;
        init_driver(s, v);
        build_desc(s, v);
        SV_TRY(hipMemcpyAsync(s->arena.at(v.off + v.o.hist), &kEmptyHist, sizeof(SlotHist),
                              hipMemcpyHostToDevice, s->stream));
        s->dirty = true;
        status[i] = -1;
      }
    }
    for (int i = n - 1; i >= 0; --i)
      if (status[i] >= 0) {
        rc = retire(s, (size_t)i, status[i]);
        if (rc) return rc;
      }
  }
  if (n_active) *n_active = (int32_t)s->active.size();
  if (n_queued) {
    std::lock_guard<std::mutex> g(s->mu);
    *n_queued = (int32_t)s->queue.size();
  }
  return SDOT_OK;
}

extern "C" int sdot_server_poll(sdot_server *s, int64_t id, int32_t *done, sdot_result *res,
                                double *psi) {
  if (!s || !done) return set_error(SDOT_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->done.find(id);
  if (it == s->done.end()) {
    *done = 0;
    if (id < 0 || id >= s->next_id) return set_error(SDOT_ERR_INVALID, "unknown request id");
    return SDOT_OK;
  }
  *done = 1;
  if (res) *res = it->second.res;
  if (psi) std::memcpy(psi, it->second.psi.data(), sizeof(double) * it->second.psi.size());
  s->done.erase(it);
  return SDOT_OK;
}

extern "C" int sdot_server_clock(sdot_server *s, double *t) {
  if (!s || !t) return set_error(SDOT_ERR_INVALID, "null argument");
  *t = s->now();
  return SDOT_OK;
}

extern "C" int sdot_server_request_bytes(sdot_server *s, int64_t n, double radius, int64_t *bytes) {
  if (!s || !bytes || n < 1) return set_error(SDOT_ERR_INVALID, "bad argument");
  // sizes only: a unit-cube target layout of n points stands in for the request
  Request r;
  r.N = n;
  r.R = (radius <= 0.0 || radius >= DBL_MAX) ? INFINITY : radius;
  std::vector<double> y((size_t)n * s->src.dim);
  for (int64_t j = 0; j < n; ++j)
    for (int d = 0; d < s->src.dim; ++d)
      y[j * s->src.dim + d] = s->src.lo[d] + (s->src.hi[d] - s->src.lo[d]) * ((j * (d + 3) + d) % 997) / 997.0;
  target_shape(s->src.dim, n, y.data(), r.L);
  fill_shape(s, r);
  *bytes = (int64_t)r.bytes;
  return SDOT_OK;
}

extern "C" int sdot_server_autotune(sdot_server *s, int32_t *pass_targets, int32_t *grad_qpoints) {
  if (!s) return set_error(SDOT_ERR_INVALID, "null server");
  {
    std::lock_guard<std::mutex> g(s->mu);
    if (!s->queue.empty() || !s->active.empty() || !s->done.empty())
      return set_error(SDOT_ERR_STATE, "autotune runs on an idle server (before the first submit)");
  }
  // synthetic batch: one solve per size class, targets spread over the source box
  const int dim = s->src.dim;
  const int64_t sizes[4] = {1024, 4096, 16384, 65536};
  std::vector<std::vector<double>> ys(4), nus(4);
  uint64_t st = 0x2545F4914F6CDD1Dull;
  auto uni = [&]() {
    st ^= st << 13;
    st ^= st >> 7;
    st ^= st << 17;
    return (double)(st >> 11) * (1.0 / 9007199254740992.0);
  };
  for (int i = 0; i < 4; ++i) {
    ys[i].resize((size_t)sizes[i] * dim);
    nus[i].assign(sizes[i], 1.0 / (double)sizes[i]);
    for (int64_t j = 0; j < sizes[i]; ++j)
      for (int d = 0; d < dim; ++d)
        ys[i][j * dim + d] = s->src.lo[d] + (s->src.hi[d] - s->src.lo[d]) * uni();
  }
  const int pts[4] = {1024, 2048, 4096, 8192};
  const int gqs[4] = {2048, 4096, 8192, 16384};
  const int prior_pt = s->pass_targets, prior_gq = s->grad_qpoints;
  // the batch must fit (4 slots, the arena) at the sizes with the most work
  // items, or autotune changes nothing
  {
    if (s->max_slots < 4) return set_error(SDOT_ERR_NOMEM, "autotune needs 4 solve slots");
    s->pass_targets = pts[0];
    s->grad_qpoints = gqs[0];
    size_t need = 0;
    int rc = SDOT_OK;
    for (int i = 0; i < 4 && rc == SDOT_OK; ++i) {
      Request r;
      r.N = sizes[i];
      r.R = INFINITY;
      target_shape(dim, r.N, ys[i].data(), r.L);
      rc = fill_shape(s, r);
      need += r.bytes;
    }
    s->pass_targets = prior_pt;
    s->grad_qpoints = prior_gq;
    if (rc) return rc;
    if (need > s->arena.size())
      return set_error(SDOT_ERR_NOMEM, "autotune batch needs " + std::to_string(need) +
                                           " device bytes, more than the server's HBM budget");
  }
  std::vector<int64_t> synthetic;
  // on an error: the prior sizes back, the synthetic solves gone from the
  // queue, the active list (their arena blocks released) and the results
  auto undo = [&](int rc) {
    s->pass_targets = prior_pt;
    s->grad_qpoints = prior_gq;
    auto ours = [&](int64_t id) {
      return std::find(synthetic.begin(), synthetic.end(), id) != synthetic.end();
    };
    std::lock_guard<std::mutex> g(s->mu);
    for (auto it = s->queue.begin(); it != s->queue.end();)
      it = ours((*it)->id) ? s->queue.erase(it) : it + 1;
    for (size_t i = 0; i < s->active.size();) {
      if (ours(s->active[i]->req->id)) {
        s->arena.release_block(s->active[i]->off, s->active[i]->o.total);
        s->active.erase(s->active.begin() + i);
        s->dirty = true;
      } else {
        ++i;
      }
    }
    for (int64_t id : synthetic) s->done.erase(id);
    return rc;
  };
  auto time_config = [&](int pt, int gq, double &ms) -> int {
    s->pass_targets = pt;
    s->grad_qpoints = gq;
    double best = INFINITY;
    for (int rep = 0; rep < 3; ++rep) {
      std::vector<int64_t> ids;
      for (int i = 0; i < 4; ++i) {
        sdot_request q{};
        q.n = sizes[i];
        q.y = ys[i].data();
        q.nu = nus[i].data();
        q.epsilon = 1e-2;
        q.tolerance = 0.0;
        q.solver_control_type = SDOT_CTRL_L1NORM;
        q.max_iterations = 0;  // one evaluation, then retired
        int64_t id = -1;
        int rc = sdot_server_submit(s, &q, &id);
        if (rc) return rc;
        ids.push_back(id);
        synthetic.push_back(id);
      }
      int32_t na = 0, nq = 0;
      int rc = sdot_server_step(s, &na, &nq);
      if (rc) return rc;
      if (na != 0 || nq != 0) return set_error(SDOT_ERR_NOMEM, "autotune batch did not fit the arena");
      if (rep > 0) best = std::min(best, (double)s->last_step_ms);
      for (int64_t id : ids) {
        int32_t d = 0;
        rc = sdot_server_poll(s, id, &d, nullptr, nullptr);
        if (rc) return rc;
      }
    }
    ms = best;
    return SDOT_OK;
  };
  int best_pt = s->pass_targets, best_gq = s->grad_qpoints;
  double best = INFINITY, ms = 0.0;
  for (int pt : pts) {
    int rc = time_config(pt, best_gq, ms);
    if (rc) return undo(rc);
    if (ms < best) {
      best = ms;
      best_pt = pt;
    }
  }
  best = INFINITY;
  for (int gq : gqs) {
    int rc = time_config(best_pt, gq, ms);
    if (rc) return undo(rc);
    if (ms < best) {
      best = ms;
      best_gq = gq;
    }
  }
  s->pass_targets = best_pt;
  s->grad_qpoints = best_gq;
  if (pass_targets) *pass_targets = best_pt;
  if (grad_qpoints) *grad_qpoints = best_gq;
  return SDOT_OK;
}
