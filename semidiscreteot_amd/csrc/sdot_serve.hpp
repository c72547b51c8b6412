// sdot_serve.hpp -- per-solve blocks of the serving path shared by the host
// scheduler (sdot_serve.cpp) and the batched kernels (sdot_serve_kernels.hip).
#pragma once

#include "sdot_internal.hpp"

namespace sdot {

constexpr int kHist = 5;  // L-BFGS history (deal.II SolverBFGS default)
enum : int { kActNone = 0, kActTrial = 1, kActIter = 2 };

// Argument blocks of one solve's evaluation.
struct SlotDesc {
  PrepArgs prep;
  P1Args p1;
  QFArgs qf;
  GArgs gr;
  GFArgs gf;
  FinArgs fin;
  int64_t nPpart;
};

// Device-resident L-BFGS history bookkeeping.
struct SlotHist {
  double rho[kHist];
  int32_t count, newest;
};

// Device vectors of one solve (caller target order).
struct SlotVecs {
  double *x, *g, *x0, *gv, *p;
  double *S[kHist], *Y[kHist];
  const double *nu;
  SlotHist *hist;
  int64_t N;
};

// Written by the host before a step.
struct SlotIn {
  int32_t action, first, has_prev, pad;
  double alpha, f, f_prev, tol;
};

// Written by the device during a step (F and C first: the final reduction
// writes them as a pair).  lost > 0: a history-seeded FP32 sweep lost range
// and the evaluation must be repeated unseeded.
struct SlotOut {
  double F, C, gp, l1, cw, relerr, g0, a1, lost;
};

struct ServeDev {
  SlotDesc *desc;
  int64_t *base;      // [5][bstride] work-list offsets: prep, pass1, qfin, grad, gfin
  int64_t bstride;
  int *nact;
  SlotVecs *vecs;
  SlotIn *in;
  SlotOut *out;
  unsigned grid_w;      // persistent grid of the one-wave work-list kernels
  unsigned grid_b;      // persistent grid of the 256-thread work-list kernels
  unsigned grid_slots;  // one workgroup per solve (upper bound on the batch)
};

hipError_t launch_serve_step(int dim, int G, int mode, const ServeDev &sd, hipStream_t s);

}  // namespace sdot
