// topology_optimizer.cpp — the reference's C++ optimizer API (topology_optimizer.hpp:120-279)
// implemented over the C-ABI of libamt.so. Every numeric phase runs on the GPU; this layer
// only converts types, uploads changed inputs and re-throws the reference's exceptions.
#include "../../include/am/topology_optimizer.hpp"

#include <cstdio>
#include <cstring>
#include <iomanip>
#include <iostream>

#include "../../include/amt.h"

namespace amt {
void hex8_stiffness(double E, double nu, double h, double* Ke);
}

namespace am {
namespace fem {

std::vector<double> elementStiffness(const MaterialProperties& mat, double element_size) {
  std::vector<double> Ke(576);
  amt::hex8_stiffness(mat.E, mat.nu, element_size, Ke.data());
  return Ke;
}

std::vector<int64_t> getElementDofIndices(int i, int j, int k, int nx, int ny, int nz) {
  (void)nx;
  static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  std::vector<int64_t> d(24);
  const int64_t ny1 = ny + 1, nz1 = nz + 1;
  for (int n = 0; n < 8; ++n) {
    const int64_t node =
        (int64_t)(i + off[n][0]) * ny1 * nz1 + (int64_t)(j + off[n][1]) * nz1 + (k + off[n][2]);
    for (int c = 0; c < 3; ++c) d[3 * n + c] = 3 * node + c;
  }
  return d;
}

}  // namespace fem

namespace topopt {

void TopologyOptimizer::check(int rc) const {
  if (rc == AMT_OK) return;
  const std::string msg = amt_last_error(h_);
  switch (rc) {
    case AMT_EINVAL:
    case AMT_ENAN:
      throw std::invalid_argument(msg);
    case AMT_ERANGE:
      throw std::out_of_range(msg);
    default:
      throw std::runtime_error(msg);
  }
}

TopologyOptimizer::TopologyOptimizer(int nx, int ny, int nz, Scalar element_size,
                                     const fem::MaterialProperties& material,
                                     const SIMPConfig& config)
    : nx_(nx), ny_(ny), nz_(nz), element_size_(element_size), config_(config) {
  n_elements_ = (int64_t)nx * ny * nz;
  n_dofs_ = 3LL * (nx + 1) * (ny + 1) * (nz + 1);
  amt_config c;
  amt_config_default(&c);
  c.nx = nx;
  c.ny = ny;
  c.nz = nz;
  c.element_size = element_size;
  c.E = material.E;
  c.nu = material.nu;
  c.volume_fraction = config.volume_fraction;
  c.penalty = config.penalty;
  c.filter_radius = config.filter_radius;
  c.move_limit = config.move_limit;
  c.rho_min = config.rho_min;
  c.E_min = config.E_min;
  c.max_iterations = config.max_iterations;
  c.convergence_tol = config.convergence_tol;
  c.pcg_tol = config.solver_tolerance;
  c.pcg_maxit = config.solver_max_iterations;
  c.penalty_factor = config.penalty_factor;
  c.density_index = (int)config.density_index;
  c.precision = (int)config.precision;
  c.precond = (int)config.preconditioner;
  // the multigrid solver starts from the energy-minimising combination of the last two
  // displacements (amt_config.warm_start 2), the other solvers from the last one
  c.warm_start = config.warm_start ? (config.preconditioner == Preconditioner::Multigrid ? 2 : 1) : 0;
  c.inner_tol = config.inner_tolerance;
  c.verbose = 0;
  c.use_graphs = 1;
  const int rc = amt_create(&c, &h_);
  if (rc != AMT_OK) {
    const std::string msg = amt_last_error(nullptr);
    if (rc == AMT_EINVAL) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
  }
  if (config_.verbose) {
    std::cout << "TopologyOptimizer initialized:" << std::endl;
    std::cout << "  Grid: " << nx_ << " x " << ny_ << " x " << nz_ << std::endl;
    std::cout << "  Elements: " << n_elements_ << std::endl;
    std::cout << "  DOFs: " << n_dofs_ << std::endl;
    std::cout << "  Filter radius: " << config_.filter_radius << std::endl;
    std::cout << "  Volume fraction: " << config_.volume_fraction << std::endl;
  }
}

TopologyOptimizer::~TopologyOptimizer() { amt_destroy(h_); }

void TopologyOptimizer::uploadLoads(const Vector& forces, const std::vector<int64_t>& fixed) {
  if (forces == loaded_forces_ && fixed == loaded_fixed_) return;
  check(amt_set_loads(h_, forces.data(), (int64_t)forces.size(), fixed.data(),
                      (int64_t)fixed.size()));
  loaded_forces_ = forces;
  loaded_fixed_ = fixed;
}

static IterationResult convert(const amt_iteration_result& r) {
  IterationResult o;
  o.compliance = r.compliance;
  o.volume_fraction = r.volume_fraction;
  o.max_density_change = r.max_density_change;
  o.solver_iterations = r.solver_iterations;
  o.assembly_time_s = r.assembly_time_s;
  o.solve_time_s = r.solve_time_s;
  o.sensitivity_time_s = r.sensitivity_time_s;
  o.update_time_s = r.update_time_s;
  o.solver_converged = r.solver_converged != 0;
  o.true_rel_residual = r.true_rel_residual;
  o.bisect_passes = r.bisect_passes;
  return o;
}

// topology_optimizer.cpp:294-384. A fresh optimizer that was never run() gets its design
// mask from this density (the reference reads an empty mask there).
IterationResult TopologyOptimizer::runIteration(std::vector<Scalar>& density,
                                                const Vector& forces,
                                                const std::vector<int64_t>& fixed_dofs) {
  uploadLoads(forces, fixed_dofs);
  check(amt_set_density(h_, density.data(), (int64_t)density.size()));
  if (!mask_built_) {
    check(amt_begin_run(h_));
    mask_built_ = true;
  }
  amt_iteration_result r;
  check(amt_run_iteration(h_, &r));
  check(amt_get_density(h_, density.data(), (int64_t)density.size()));
  if (!r.solver_converged) std::cerr << "WARNING: Solver did not converge at iteration!" << std::endl;
  return convert(r);
}

namespace {
struct CbCtx {
  IterationCallback* cb;
  std::vector<Scalar> buf;
};
int cb_thunk(int it, const double* d, int64_t n, const amt_iteration_result* r, void* user) {
  CbCtx* c = (CbCtx*)user;
  c->buf.assign(d, d + n);
  return (*c->cb)(it, c->buf, convert(*r)) ? 1 : 0;
}
}  // namespace

// topology_optimizer.cpp:391-520
OptimizationResult TopologyOptimizer::run(std::vector<Scalar> initial_density,
                                          const Vector& forces,
                                          const std::vector<int64_t>& fixed_dofs,
                                          IterationCallback callback) {
  OptimizationResult result;
  if ((int64_t)initial_density.size() != n_elements_)  // :403-405
    initial_density.assign(n_elements_, config_.volume_fraction);
  uploadLoads(forces, fixed_dofs);
  check(amt_set_density(h_, initial_density.data(), n_elements_));
  std::vector<amt_iteration_result> hist(std::max(config_.max_iterations, 0));
  int iters = 0, conv = 0;
  double elapsed = 0.0;
  CbCtx ctx{&callback, {}};
  check(amt_run(h_, config_.max_iterations, hist.data(), &iters, &conv, &elapsed,
                callback ? cb_thunk : nullptr, callback ? &ctx : nullptr));
  mask_built_ = true;
  for (int i = 0; i < iters; ++i) {
    result.compliance_history.push_back(hist[i].compliance);
    result.volume_history.push_back(hist[i].volume_fraction);
    if (config_.verbose)
      std::cout << std::setw(6) << i << std::setw(14) << std::scientific << std::setprecision(4)
                << hist[i].compliance << std::setw(10) << std::fixed << std::setprecision(3)
                << hist[i].volume_fraction << std::setw(12) << std::scientific
                << std::setprecision(2) << hist[i].max_density_change << std::setw(10)
                << hist[i].solver_iterations << std::endl;
  }
  result.iterations = iters;
  result.converged = conv != 0;
  result.elapsed_time_s = elapsed;
  result.final_density.resize(n_elements_);
  check(amt_get_density(h_, result.final_density.data(), n_elements_));
  return result;
}

std::vector<FilterWeight> TopologyOptimizer::filterNeighbors(int64_t e) const {
  int count = 0;
  check(amt_filter_neighbors(h_, e, nullptr, nullptr, 0, &count));
  std::vector<int64_t> idx(count);
  std::vector<double> w(count);
  check(amt_filter_neighbors(h_, e, idx.data(), w.data(), count, &count));
  std::vector<FilterWeight> out(count);
  for (int i = 0; i < count; ++i) out[i] = {idx[i], w[i]};
  return out;
}

Vector TopologyOptimizer::displacement() const {
  Vector u(n_dofs_);
  check(amt_get_displacement(h_, u.data(), n_dofs_));
  return u;
}

}  // namespace topopt
}  // namespace am
