// sdot_sot_solver.hpp -- header-only C++ facade with the method set of the
// reference's SotSolver<dim, spacedim> (include/SemiDiscreteOT/solvers/SotSolver.h:
// 176-348) over the C ABI of sdot.h.  No deal.II types: setup_source takes the
// flat arrays a deal.II caller pulls once from FEValues (one loop over the
// locally owned cells: quadrature points, density times JxW, cell centers).
// Errors throw std::runtime_error as the reference does for invalid measures
// (SotSolver.cc:110-112); non-convergence does not throw (SotSolver.cc:207-211)
// and is reported by get_convergence_status().
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "sdot.h"

namespace sdot {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string &msg) : std::runtime_error(msg), code(code) {}
  int code;
};

inline void check(int rc) {
  if (rc != SDOT_OK) throw Error(rc, sdot_last_error());
}

template <int dim, int spacedim = dim>
class SotSolverGPU {
  static_assert(spacedim == 2 || spacedim == 3, "spacedim must be 2 or 3");

 public:
  using Params = sdot_params;

  explicit SotSolverGPU(int device = 0) {
    check(sdot_ctx_create(device, &ctx_));
    check(sdot_params_default(&params_));
  }
  ~SotSolverGPU() { sdot_ctx_destroy(ctx_); }
  SotSolverGPU(const SotSolverGPU &) = delete;
  SotSolverGPU &operator=(const SotSolverGPU &) = delete;

  static Params default_parameters() {
    Params p;
    check(sdot_params_default(&p));
    return p;
  }
  static Params parameters_from_prm(const std::string &path) {
    Params p = default_parameters();
    check(sdot_params_load_prm(path.c_str(), &p));
    return p;
  }

  // SotSolver::setup_source: x [n_cells*nq*spacedim], w = rho*JxW [n_cells*nq],
  // centers [n_cells*spacedim].
  void setup_source(const std::vector<double> &x, const std::vector<double> &w,
                    const std::vector<double> &centers, int nq) {
    const int64_t n_cells = (int64_t)(centers.size() / spacedim);
    if ((int64_t)w.size() != n_cells * nq || (int64_t)x.size() != n_cells * nq * spacedim)
      throw Error(SDOT_ERR_INVALID, "setup_source: inconsistent array sizes");
    check(sdot_set_source(ctx_, spacedim, n_cells, nq, x.data(), w.data(), centers.data()));
  }

  // SotSolver::setup_target: points [N*spacedim], weights [N].
  void setup_target(const std::vector<double> &points, const std::vector<double> &weights) {
    if (points.size() != weights.size() * spacedim)
      throw Error(SDOT_ERR_INVALID, "setup_target: dimension mismatch");
    n_ = (int64_t)weights.size();
    check(sdot_set_target(ctx_, n_, points.data(), weights.data()));
  }

  void configure(const Params &p) {
    params_ = p;
    check(sdot_set_params(ctx_, &params_));
  }

  // SotSolver::solve(potential, params): potential resized to N if needed
  // (SotSolver.cc:147-150).
  void solve(std::vector<double> &potential, const Params &p) {
    configure(p);
    if (n_ == 0) throw Error(SDOT_ERR_STATE, "Invalid measures configuration");
    if ((int64_t)potential.size() != n_) potential.assign(n_, 0.0);
    check(sdot_solve(ctx_, potential.data(), &info_));
  }

  // SotSolver::evaluate_functional
  double evaluate_functional(const std::vector<double> &potential, std::vector<double> &gradient_out) {
    gradient_out.resize(n_);
    double F = 0, C = 0;
    check(sdot_eval(ctx_, potential.data(), gradient_out.data(), &F, &C));
    return F;
  }

  // SotSolver::compute_hessian (row-major N x N)
  void compute_hessian(const std::vector<double> &potential, std::vector<double> &hessian_out) {
    hessian_out.resize((size_t)n_ * n_);
    check(sdot_hessian(ctx_, potential.data(), hessian_out.data(), n_));
  }

  // SotSolver::evaluate_weighted_barycenters (SotSolver.cc:585-610): [N*spacedim].
  void evaluate_weighted_barycenters(const std::vector<double> &potential,
                                     std::vector<double> &barycenters_out) {
    barycenters_out.resize((size_t)n_ * spacedim);
    check(sdot_barycenters(ctx_, potential.data(), barycenters_out.data()));
  }

  // SotSolver::get_potential_conditioned_density (SotSolver.cc:973-1046) at the
  // caller's support points [n*spacedim] with the density there [n]; row k of
  // the result [indices.size()*n] belongs to potential_indices[k].  Threshold:
  // the last distance threshold, epsilon: the configured one.
  void get_potential_conditioned_density(const std::vector<double> &points,
                                         const std::vector<double> &density,
                                         const std::vector<double> &potential,
                                         const std::vector<int64_t> &potential_indices,
                                         std::vector<double> &conditioned_densities) {
    const int64_t n = (int64_t)density.size();
    if ((int64_t)points.size() != n * spacedim)
      throw Error(SDOT_ERR_INVALID, "conditioned density: points and density differ in size");
    conditioned_densities.assign(potential_indices.size() * (size_t)n, 0.0);
    check(sdot_conditional_density(ctx_, potential.data(), get_last_distance_threshold(), n,
                                   points.data(), density.data(),
                                   (int64_t)potential_indices.size(), potential_indices.data(),
                                   conditioned_densities.data()));
  }

  double get_last_functional_value() const { return state().last_functional; }
  unsigned int get_last_iteration_count() const { return (unsigned)state().last_iteration_count; }
  bool get_convergence_status() const { return state().converged != 0; }
  double get_last_distance_threshold() const { return state().last_distance_threshold; }
  double get_C_global() const { return state().C_global; }
  void set_distance_threshold(double t) { check(sdot_set_distance_threshold(ctx_, t)); }
  double compute_covering_radius() const {
    double r = 0;
    check(sdot_covering_radius(ctx_, &r));
    return r;
  }
  const sdot_solve_info &last_solve_info() const { return info_; }
  sdot_ctx *handle() const { return ctx_; }

 private:
  sdot_state state() const {
    sdot_state s;
    check(sdot_get_state(ctx_, &s));
    return s;
  }
  sdot_ctx *ctx_ = nullptr;
  Params params_{};
  sdot_solve_info info_{};
  int64_t n_ = 0;
};

// SoftmaxRefinement<dim, spacedim> (SoftmaxRefinement.h, SoftmaxRefinement.cc:4-27,
// 200-280) over the source measure of a SotSolverGPU: the reference's constructor
// takes the source DoFHandler/mapping/FE/density/quadrature order, which the
// solver's context already holds as quadrature data.
template <int dim, int spacedim = dim>
class SoftmaxRefinementGPU {
 public:
  SoftmaxRefinementGPU(const SotSolverGPU<dim, spacedim> &solver, double distance_threshold,
                       bool use_log_sum_exp_trick)
      : ctx_(solver.handle()), threshold_(distance_threshold), lse_(use_log_sum_exp_trick) {}

  // Points flat (n*spacedim); child_indices[level][c] = fine children of coarse c.
  // target_density_fine is not read by the reference's algorithm.
  std::vector<double> compute_refinement(
      const std::vector<double> &target_points_fine, const std::vector<double> &target_density_fine,
      const std::vector<double> &target_points_coarse,
      const std::vector<double> &target_density_coarse, const std::vector<double> &potential_coarse,
      double regularization_param, int level,
      const std::vector<std::vector<std::vector<size_t>>> &child_indices) const {
    (void)target_density_fine;
    const int64_t n_fine = (int64_t)(target_points_fine.size() / spacedim);
    std::vector<double> out(n_fine, 0.0);
    if (level < 0 || level >= (int)child_indices.size()) return out;  // reference: error, zeros
    const auto &lists = child_indices[level];
    std::vector<int64_t> off(lists.size() + 1, 0), flat;
    for (size_t c = 0; c < lists.size(); ++c) {
      for (size_t f : lists[c]) flat.push_back(f > (size_t)INT64_MAX ? -1 : (int64_t)f);
      off[c + 1] = (int64_t)flat.size();
    }
    check(sdot_softmax_refinement(ctx_, n_fine, target_points_fine.data(),
                                  (int64_t)target_density_coarse.size(),
                                  target_points_coarse.data(), target_density_coarse.data(),
                                  potential_coarse.data(), (int64_t)lists.size(), off.data(),
                                  flat.empty() ? nullptr : flat.data(), regularization_param,
                                  threshold_, lse_ ? 1 : 0, out.empty() ? nullptr : out.data()));
    return out;
  }

 private:
  sdot_ctx *ctx_;
  double threshold_;
  bool lse_;
};

// OptimalTransportPlan<spacedim> (OptimalTransportPlan.h, OptimalTransportPlan.cc:
// 8-96) with its "modal" and "barycentric" strategies (:99-189, :218-310).
// Points flat (n*spacedim).  The source points are held by a context as
// one-point cells; the map reads only the targets.
template <int spacedim>
class OptimalTransportPlanGPU {
  static_assert(spacedim == 2 || spacedim == 3, "spacedim must be 2 or 3");

 public:
  explicit OptimalTransportPlanGPU(const std::string &strategy_name = "modal", int device = 0)
      : device_(device) {
    set_strategy(strategy_name);
  }
  void set_strategy(const std::string &name) {
    if (name == "modal")
      strategy_ = SDOT_MAP_MODAL;
    else if (name == "barycentric")
      strategy_ = SDOT_MAP_BARYCENTRIC;
    else
      throw std::runtime_error("Unknown strategy: " + name);
  }
  static std::vector<std::string> get_available_strategies() { return {"modal", "barycentric"}; }
  void set_source_measure(const std::vector<double> &points, const std::vector<double> &density) {
    if (points.size() != density.size() * spacedim)
      throw Error(SDOT_ERR_INVALID, "source points and density differ in size");
    src_ = points;
    rho_ = density;
  }
  void set_target_measure(const std::vector<double> &points, const std::vector<double> &density) {
    if (points.size() != density.size() * spacedim)
      throw Error(SDOT_ERR_INVALID, "target points and density differ in size");
    tgt_ = points;
    nu_ = density;
  }
  void set_potential(const std::vector<double> &potential, double regularization_param) {
    if (potential.size() != nu_.size())
      throw Error(SDOT_ERR_INVALID, "potential size differs from the target count");
    psi_ = potential;
    eps_ = regularization_param;
  }
  void set_truncation_radius(double radius) { radius_ = radius; }
  void compute_map() {
    if (rho_.empty()) throw std::runtime_error("Source measure not set");
    if (nu_.empty()) throw std::runtime_error("Target measure not set");
    if (psi_.empty()) throw std::runtime_error("Transport potential not set");
    SotSolverGPU<spacedim, spacedim> s(device_);
    std::vector<double> w(rho_.size());
    for (size_t i = 0; i < w.size(); ++i) w[i] = rho_[i] < 0 ? -rho_[i] : rho_[i];
    s.setup_source(src_, w, src_, 1);
    s.setup_target(tgt_, nu_);
    mapped_points.assign(src_.size(), 0.0);
    check(sdot_transport_map(s.handle(), strategy_, psi_.data(), eps_, radius_,
                             (int64_t)rho_.size(), src_.data(), mapped_points.data()));
    transport_density = rho_;
  }
  std::vector<double> mapped_points, transport_density;

 private:
  int device_, strategy_ = SDOT_MAP_MODAL;
  std::vector<double> src_, rho_, tgt_, nu_, psi_;
  double eps_ = 1e-2, radius_ = -1.0;
};

}  // namespace sdot
