// problem.hpp — the optimizer's inputs as the reference describes them: a voxel domain
// (src/am/core/geometry.py:16-76, int8 grid: -1 void, 0 design, 1 fixed) and a load case
// (src/am/core/loads.py:107-156: point loads and constrained DOFs), plus the BRK-A-01 bracket
// set-up of the reference's C++ driver (cpp_engine/src/am_topopt.cpp:44-334, 561-572).
//
// Index conventions (the reference's): the domain grid and the density are in the
// optimizer's x-fastest slot order, element (i, j, k) at k*ny*nx + j*nx + i
// (am_topopt.cpp BracketDomain::index); nodes are z-fastest, node (i, j, k) at
// (i*(ny+1) + j)*(nz+1) + k, DOF = 3*node + component (element_kernel.cpp:186-216).
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace am {

enum class ElementType : int8_t { Void = -1, Design = 0, Fixed = 1 };

struct VoxelDomain {
  int nx = 0, ny = 0, nz = 0;
  std::vector<int8_t> grid;  // slot order (x fastest)
  double resolution_mm = 1.0;
  std::array<double, 3> origin{{0.0, 0.0, 0.0}};

  int64_t index(int i, int j, int k) const {
    return (int64_t)k * ny * nx + (int64_t)j * nx + i;
  }
  int64_t nElements() const { return (int64_t)nx * ny * nz; }
  int64_t count(ElementType t) const;
  // am_topopt.cpp:561-572: design = volume fraction, fixed = 1, void = void_value (slot order)
  std::vector<double> initialDensity(double volume_fraction, double void_value) const;
};

inline int64_t nodeId(int i, int j, int k, int ny, int nz) {
  return ((int64_t)i * (ny + 1) + j) * (nz + 1) + k;
}

struct LoadCase {
  std::string name;
  int64_t n_dofs = 0;
  std::vector<std::pair<int64_t, std::array<double, 3>>> point_loads;  // node, force
  std::vector<int64_t> constrained;                                    // DOFs
  std::vector<double> getForceVector() const;       // sum of the point loads per DOF
  std::vector<int64_t> getConstrainedDofs() const;  // sorted, unique
};

// am_topopt.cpp:44-77 (geometry and load of the BRK-A-01 bracket, millimetres / newtons)
struct BracketConfig {
  double size_x_mm = 120.0, size_y_mm = 60.0, size_z_mm = 80.0;
  double resolution_mm = 2.0;
  double base_thickness_mm = 5.0;
  double hole_diameter_mm = 8.0, hole_offset_mm = 15.0;
  double eyelet_diameter_mm = 12.0, eyelet_offset_mm = 4.0;
  double eyelet_z_offset_mm = 15.0, eyelet_height_mm = 20.0;
  double load_force_N = 15000.0;
};

// BracketDomain (am_topopt.cpp:98-239): fixed base slab with four void bolt holes, a fixed
// eyelet ring with a void bore near the top, and a fixed pillar joining them.
VoxelDomain bracketDomain(const BracketConfig& cfg);
// createForceVector + createFixedDofs (am_topopt.cpp:248-334): the load spread over the
// eyelet's ring nodes at mid-height along -z; every DOF of the z = 0 nodes clamped.
LoadCase bracketLoadCase(const VoxelDomain& d, const BracketConfig& cfg);

}  // namespace am
