// sdot_serve.hpp -- per-solve blocks of the serving path shared by the host
// scheduler (sdot_serve.cpp) and the batched kernels (sdot_serve_kernels.hip).
#pragma once

#include "sdot_internal.hpp"
#include "sdot_serve_driver.hpp"

namespace sdot {

// Reductions of one solve's value bound (VBound, sdot_device.hpp): the largest
// psi + eps ln nu per target bin and row, the smallest V_q per cell, cell bin
// and row.
struct VBArgs {
  const double *psit_r;
  const int32_t *tstart;
  double *pbin, *prow;
  int64_t ntb, ntr;
  int tnx;
  const double *qrec_r;
  const int32_t *cstart;
  double *vcell, *vbin, *vrow;
  int64_t n_cells, ncb, ncr;
  int cnx, nq, log2_units;
  double theta, inv_coef;
};

// Argument blocks of one solve's evaluation.
struct SlotDesc {
  PrepArgs prep;
  P1Args p1;
  QFArgs qf;
  GArgs gr;
  GFArgs gf;
  FinArgs fin;
  VBArgs vb;
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

// Work lists of one step.  Kernels that do not depend on the arithmetic mode
// run over every solve; the mode-templated passes run over the solves of their
// mode (the other solves contribute no items).
// FP32 solves whose evaluations are value-bounded run their own sweep lists
// (V) and the bound's reductions (blocks of 256 bins / rows / cells).
enum : int {
  kLPrep = 0,
  kLPass32, kLPass32V, kLPass64,
  kLQfin32, kLQfin64,
  kLGrad32, kLGrad32V, kLGrad64,
  kLGfin,
  kLTbin, kLTrow, kLCell, kLCbin, kLCrow,
  kNumLists
};

// Item counters of the dynamically scheduled sweeps (one per kernel and XCD).
enum : int {
  kQPass2F32 = 0, kQPass2F32V, kQPass0F64, kQPass1F64, kQGradF32, kQGradF32V, kQGradF64,
  kNumQueues
};

struct ServeDev {
  SlotDesc *desc;
  int *ctr;           // [kNumQueues][8], zeroed by the first kernel of every step
  int64_t *base;      // [kNumLists][bstride] item offsets per solve
  int64_t bstride;
  int *nact;
  SlotVecs *vecs;
  SlotIn *in;
  SlotOut *out;
  // persistent grids: as many workgroups as the device holds at once
  unsigned grid_pass;   // q-major sweeps (one wave each)
  unsigned grid_grad;   // target-major sweeps (one wave each)
  unsigned grid_b;      // 256-thread work-list kernels
  unsigned grid_slots;  // one workgroup per solve (upper bound on the batch)
};

// G: points per lane of the unbounded q-major sweeps; Gs (at most 4): of the
// value-bounded ones.
hipError_t launch_serve_step(int dim, int G, int Gs, const ServeDev &sd, hipStream_t s);
// Sets the persistent grid sizes from the occupancy of the step's kernels.
hipError_t serve_grid_sizes(int dim, int G, int Gs, int device, ServeDev &sd);

}  // namespace sdot
