// host_util.hpp — host helpers shared by the single-problem and batched engines.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "device.hpp"

namespace amt {

// Filter stencil (topology_optimizer.cpp:60-102): offsets in dz, dy, dx order with the
// reference's distance and weight arithmetic; the per-element list is this table with
// out-of-grid offsets skipped.
inline void build_stencil(double r, std::vector<StencilEntry>& st, std::vector<double>& w) {
  const int R = (int)std::ceil(r);
  for (int dz = -R; dz <= R; ++dz)
    for (int dy = -R; dy <= R; ++dy)
      for (int dx = -R; dx <= R; ++dx) {
        const double dist = std::sqrt((double)(dx * dx + dy * dy + dz * dz));
        if (dist <= r) {
          st.push_back({dx, dy, dz, 0});
          w.push_back(std::max(0.0, r - dist));
        }
      }
}

inline double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace amt
