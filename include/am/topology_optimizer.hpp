/**
 * @file topology_optimizer.hpp
 * @brief Drop-in C++ API of the SIMP optimizer, running on one MI355X through libamt.so.
 *
 * Same names, argument meaning and error behaviour as the reference's
 * cpp_engine/include/am/topology_optimizer.hpp:49-279 and element_kernel.hpp:61-87, with
 * two differences forced by the build: Eigen is absent, so `Vector` is std::vector<double>;
 * and SIMPConfig carries the solver settings the reference hard-codes in its constructor
 * (topology_optimizer.cpp:35-40) plus the device options (precision, density index).
 */
#pragma once

#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

struct amt_problem;

namespace am {
namespace fem {

/// element_kernel.hpp:61-87 (Ti6Al4V defaults)
struct MaterialProperties {
  double E;
  double nu;
  double rho;
  static constexpr double DEFAULT_E = 113.8e9;
  static constexpr double DEFAULT_NU = 0.342;
  static constexpr double DEFAULT_RHO = 4430.0;
  MaterialProperties(double E_ = DEFAULT_E, double nu_ = DEFAULT_NU, double rho_ = DEFAULT_RHO)
      : E(E_), nu(nu_), rho(rho_) {}
  double lambda() const { return E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)); }
  double mu() const { return E / (2.0 * (1.0 + nu)); }
};

/// Hex8 Ke0 (24x24 row-major) of a cube of edge h (element_kernel.cpp:155-179).
std::vector<double> elementStiffness(const MaterialProperties& mat, double element_size);

/// 24 global DOFs of element (i,j,k) (element_kernel.cpp:186-216).
std::vector<int64_t> getElementDofIndices(int i, int j, int k, int nx, int ny, int nz);

}  // namespace fem

namespace topopt {

using Scalar = double;
using Vector = std::vector<double>;

enum class DensityIndex { Consistent = 0, Reference = 1 };
enum class Precision { FP64 = 0, MixedFP32 = 1, MixedBF16 = 2 };
/// Jacobi: the reference's CG with Jacobi in place of IncompleteCholesky. Multigrid: FP64
/// flexible CG around an FP32 Galerkin geometric-multigrid V-cycle.
enum class Preconditioner { Jacobi = 0, Multigrid = 2 };

/// topology_optimizer.hpp:49-59, plus the solver settings of topology_optimizer.cpp:35-40
struct SIMPConfig {
  Scalar volume_fraction = 0.25;
  Scalar penalty = 3.0;
  Scalar filter_radius = 2.0;
  Scalar move_limit = 0.2;
  Scalar rho_min = 0.01;
  Scalar E_min = 1e-6;
  int max_iterations = 100;
  Scalar convergence_tol = 0.01;
  bool verbose = true;
  // solver (reference: fixed inside the constructor)
  Scalar solver_tolerance = 1e-6;
  int solver_max_iterations = 5000;
  Scalar penalty_factor = 1e8;
  // device options
  DensityIndex density_index = DensityIndex::Reference;
  Precision precision = Precision::FP64;
  Preconditioner preconditioner = Preconditioner::Jacobi;
  bool warm_start = false;  // Multigrid: from the energy-minimising combination of the last two
                            // displacements; Jacobi: from the last displacement
  Scalar inner_tolerance = 1e-4;
};

/// topology_optimizer.hpp:65-74
struct IterationResult {
  double compliance = 0.0;
  double volume_fraction = 0.0;
  double max_density_change = 0.0;
  int solver_iterations = 0;
  double assembly_time_s = 0.0;
  double solve_time_s = 0.0;
  double sensitivity_time_s = 0.0;
  double update_time_s = 0.0;
  // solver detail (not in the reference struct)
  bool solver_converged = false;
  double true_rel_residual = 0.0;
  int bisect_passes = 0;
};

/// topology_optimizer.hpp:80-87
struct OptimizationResult {
  std::vector<Scalar> final_density;
  std::vector<double> compliance_history;
  std::vector<double> volume_history;
  int iterations = 0;
  bool converged = false;
  double elapsed_time_s = 0.0;
};

/// topology_optimizer.hpp:98-102
using IterationCallback =
    std::function<bool(int iteration, const std::vector<Scalar>& density,
                       const IterationResult& result)>;

/// topology_optimizer.hpp:108-111
struct FilterWeight {
  int64_t neighbor_idx;
  Scalar weight;
};

/// topology_optimizer.hpp:120-279
class TopologyOptimizer {
 public:
  TopologyOptimizer(int nx, int ny, int nz, Scalar element_size,
                    const fem::MaterialProperties& material,
                    const SIMPConfig& config = SIMPConfig());
  ~TopologyOptimizer();
  TopologyOptimizer(const TopologyOptimizer&) = delete;
  TopologyOptimizer& operator=(const TopologyOptimizer&) = delete;

  IterationResult runIteration(std::vector<Scalar>& density, const Vector& forces,
                               const std::vector<int64_t>& fixed_dofs);

  OptimizationResult run(std::vector<Scalar> initial_density, const Vector& forces,
                         const std::vector<int64_t>& fixed_dofs,
                         IterationCallback callback = nullptr);

  int nx() const { return nx_; }
  int ny() const { return ny_; }
  int nz() const { return nz_; }
  int64_t numElements() const { return n_elements_; }
  int64_t numDofs() const { return n_dofs_; }
  Scalar elementSize() const { return element_size_; }
  const SIMPConfig& config() const { return config_; }

  /// Neighbour list of element slot e as the device filter walks it.
  std::vector<FilterWeight> filterNeighbors(int64_t e) const;
  /// Displacement of the last solve.
  Vector displacement() const;

 private:
  int nx_, ny_, nz_;
  int64_t n_elements_, n_dofs_;
  Scalar element_size_;
  SIMPConfig config_;
  amt_problem* h_ = nullptr;
  bool mask_built_ = false;
  Vector loaded_forces_;
  std::vector<int64_t> loaded_fixed_;

  void uploadLoads(const Vector& forces, const std::vector<int64_t>& fixed_dofs);
  void check(int rc) const;
};

}  // namespace topopt
}  // namespace am
