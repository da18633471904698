// problem.cpp — VoxelDomain / LoadCase and the BRK-A-01 bracket set-up (include/am/problem.hpp).
// The bracket rules follow cpp_engine/src/am_topopt.cpp:98-334 step for step (grid sizes by
// truncation, cylinder bounding boxes, the ring-load radii and fallback): the set-up is part
// of the reference's on-disk results (data/cpp_test_run), so it has to be the same.
#include "../../include/am/problem.hpp"

#include <algorithm>
#include <cmath>

namespace am {

int64_t VoxelDomain::count(ElementType t) const {
  return std::count(grid.begin(), grid.end(), static_cast<int8_t>(t));
}

std::vector<double> VoxelDomain::initialDensity(double volume_fraction, double void_value) const {
  std::vector<double> rho(grid.size());
  for (size_t e = 0; e < grid.size(); ++e) {
    const ElementType t = static_cast<ElementType>(grid[e]);
    rho[e] = t == ElementType::Design ? volume_fraction
                                      : (t == ElementType::Fixed ? 1.0 : void_value);
  }
  return rho;
}

std::vector<double> LoadCase::getForceVector() const {
  std::vector<double> F(n_dofs, 0.0);
  for (const auto& pl : point_loads)
    for (int c = 0; c < 3; ++c) F[3 * pl.first + c] += pl.second[c];
  return F;
}

std::vector<int64_t> LoadCase::getConstrainedDofs() const {
  std::vector<int64_t> d = constrained;
  std::sort(d.begin(), d.end());
  d.erase(std::unique(d.begin(), d.end()), d.end());
  return d;
}

namespace {

// Elements whose column (i, j) lies within radius of (cx, cy), layers [z0, z0 + height).
// A fixed cylinder leaves void elements void (markCylinderFixed); a void one overwrites.
void mark_cylinder(VoxelDomain& d, double cx, double cy, int z0, double radius, int height,
                   ElementType val) {
  const int i0 = std::max(0, (int)(cx - radius - 1)), i1 = std::min(d.nx, (int)(cx + radius + 2));
  const int j0 = std::max(0, (int)(cy - radius - 1)), j1 = std::min(d.ny, (int)(cy + radius + 2));
  for (int i = i0; i < i1; ++i)
    for (int j = j0; j < j1; ++j) {
      const double di = i - cx, dj = j - cy;
      if (di * di + dj * dj > radius * radius) continue;
      for (int k = z0; k < std::min(d.nz, z0 + height); ++k) {
        int8_t& e = d.grid[d.index(i, j, k)];
        if (val == ElementType::Void || e != static_cast<int8_t>(ElementType::Void))
          e = static_cast<int8_t>(val);
      }
    }
}

struct Eyelet {
  int z0, z1, x, y;
};

Eyelet eyelet_span(const BracketConfig& c, int nx, int ny, int nz) {
  const int zoff = (int)(c.eyelet_z_offset_mm / c.resolution_mm);
  const int h = std::max(1, (int)(c.eyelet_height_mm / c.resolution_mm));
  const int zc = nz - zoff;
  Eyelet e;
  e.z0 = std::max(0, zc - h / 2);
  e.z1 = std::min(nz, e.z0 + h);
  e.x = nx / 2;
  e.y = ny / 2;
  return e;
}

}  // namespace

VoxelDomain bracketDomain(const BracketConfig& c) {
  VoxelDomain d;
  d.resolution_mm = c.resolution_mm;
  d.nx = (int)(c.size_x_mm / c.resolution_mm);
  d.ny = (int)(c.size_y_mm / c.resolution_mm);
  d.nz = (int)(c.size_z_mm / c.resolution_mm);
  d.grid.assign((size_t)d.nx * d.ny * d.nz, static_cast<int8_t>(ElementType::Design));
  // base slab
  const int base_z = std::max(1, (int)(c.base_thickness_mm / c.resolution_mm));
  for (int k = 0; k < std::min(base_z, d.nz); ++k)
    for (int j = 0; j < d.ny; ++j)
      for (int i = 0; i < d.nx; ++i)
        d.grid[d.index(i, j, k)] = static_cast<int8_t>(ElementType::Fixed);
  // four bolt holes through the base
  const double hole_r = c.hole_diameter_mm / 2.0 / c.resolution_mm;
  const int hox = (int)(c.hole_offset_mm / c.resolution_mm);
  const int hoy = (int)(c.hole_offset_mm / c.resolution_mm);
  const int holes[4][2] = {{hox, hoy}, {d.nx - hox, hoy}, {hox, d.ny - hoy}, {d.nx - hox, d.ny - hoy}};
  for (const auto& h : holes) mark_cylinder(d, h[0], h[1], 0, hole_r, base_z, ElementType::Void);
  // eyelet: fixed ring with a void bore
  const Eyelet e = eyelet_span(c, d.nx, d.ny, d.nz);
  const double ring_r = (c.eyelet_diameter_mm / 2.0 + c.eyelet_offset_mm) / c.resolution_mm;
  if (e.z1 - e.z0 > 0) {
    mark_cylinder(d, e.x, e.y, e.z0, ring_r, e.z1 - e.z0, ElementType::Fixed);
    mark_cylinder(d, e.x, e.y, e.z0, c.eyelet_diameter_mm / 2.0 / c.resolution_mm, e.z1 - e.z0,
                  ElementType::Void);
  }
  // pillar joining base and eyelet
  if (e.z0 > base_z)
    mark_cylinder(d, e.x, e.y, base_z, ring_r, e.z0 - base_z, ElementType::Fixed);
  return d;
}

LoadCase bracketLoadCase(const VoxelDomain& d, const BracketConfig& c) {
  LoadCase lc;
  lc.name = "BRK-A-01_Static_1";
  lc.n_dofs = 3LL * (d.nx + 1) * (d.ny + 1) * (d.nz + 1);
  const Eyelet e = eyelet_span(c, d.nx, d.ny, d.nz);
  const int ez = std::max(1, std::min(d.nz - 1, (e.z0 + e.z1) / 2));
  double inner_mm = c.eyelet_diameter_mm / 2.0 + c.resolution_mm * 1.25;
  const double outer_mm = c.eyelet_diameter_mm / 2.0 + c.eyelet_offset_mm;
  if (inner_mm >= outer_mm) inner_mm = c.eyelet_diameter_mm / 2.0 + c.eyelet_offset_mm * 0.5;
  const double ri = inner_mm / c.resolution_mm, ro = outer_mm / c.resolution_mm;
  const int search = (int)ro + 2;
  std::vector<int64_t> nodes;
  for (int i = std::max(0, e.x - search); i <= std::min(d.nx, e.x + search); ++i)
    for (int j = std::max(0, e.y - search); j <= std::min(d.ny, e.y + search); ++j) {
      const double dx = i - e.x, dy = j - e.y;
      const double dist = std::sqrt(dx * dx + dy * dy);
      if (dist > ri && dist <= ro) nodes.push_back(nodeId(i, j, ez, d.ny, d.nz));
    }
  if (nodes.empty()) nodes.push_back(nodeId(e.x, e.y, ez, d.ny, d.nz));
  const double fpn = c.load_force_N / (double)nodes.size();
  for (int64_t n : nodes) lc.point_loads.push_back({n, {{0.0, 0.0, -fpn}}});
  for (int i = 0; i <= d.nx; ++i)
    for (int j = 0; j <= d.ny; ++j)
      for (int comp = 0; comp < 3; ++comp) lc.constrained.push_back(3 * nodeId(i, j, 0, d.ny, d.nz) + comp);
  return lc;
}

}  // namespace am
