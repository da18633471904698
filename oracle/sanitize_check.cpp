// sanitize_check.cpp — test infrastructure: runs the CPU oracle and the product's host-only C++
// (the BRK-A-01 set-up in am_ml_suite_amd/csrc/problem.cpp and the closed-form Ke0 in
// am_ml_suite_amd/csrc/ke0.cpp) under AddressSanitizer and UndefinedBehaviorSanitizer
// (`make -C oracle sanitize`). Each step also checks its result against the other side, so a
// silent out-of-bounds read that changes a value fails as well. Exit status 0 on success.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../include/am/problem.hpp"
#include "oracle.h"

namespace amt {
void hex8_stiffness(double E, double nu, double h, double* Ke);
}

static int fails = 0;
#define CHECK(cond, ...)                          \
  do {                                            \
    if (!(cond)) {                                \
      std::fprintf(stderr, "FAIL: " __VA_ARGS__); \
      std::fprintf(stderr, "\n");                 \
      ++fails;                                    \
    }                                             \
  } while (0)

int main() {
  // closed-form Ke0 (product) against the quadrature (oracle)
  for (double nu : {0.3, 0.342, 0.1}) {
    double a[576], b[576];
    amt::hex8_stiffness(2.0, nu, 0.5, a);
    orc_ke0(2.0, nu, 0.5, b);
    double d = 0.0, m = 0.0;
    for (int i = 0; i < 576; ++i) {
      d = std::fmax(d, std::fabs(a[i] - b[i]));
      m = std::fmax(m, std::fabs(b[i]));
    }
    CHECK(d <= 1e-14 * m, "Ke0 nu=%g differs by %g", nu, d);
  }
  // BRK-A-01 set-up (product) against the oracle's restatement, at 12 mm and 4 mm
  for (double res : {12.0, 4.0}) {
    am::BracketConfig bc;
    bc.resolution_mm = res;
    const am::VoxelDomain dom = am::bracketDomain(bc);
    int nx, ny, nz;
    const int64_t ne = orc_bracket_domain(res, nullptr, &nx, &ny, &nz);
    CHECK(ne == dom.nElements() && nx == dom.nx && ny == dom.ny && nz == dom.nz,
          "bracket %g mm: grid size", res);
    std::vector<int8_t> g(ne);
    orc_bracket_domain(res, g.data(), &nx, &ny, &nz);
    CHECK(g == dom.grid, "bracket %g mm: grid", res);
    const am::LoadCase lc = am::bracketLoadCase(dom, bc);
    std::vector<double> F(3 * (nx + 1) * (ny + 1) * (nz + 1));
    orc_bracket_forces(nx, ny, nz, res, bc.load_force_N, F.data());
    const std::vector<double> Fp = lc.getForceVector();
    CHECK(Fp.size() == F.size(), "bracket %g mm: force size", res);
    double df = 0.0;
    for (size_t i = 0; i < F.size() && i < Fp.size(); ++i) df = std::fmax(df, std::fabs(F[i] - Fp[i]));
    CHECK(df <= 1e-9, "bracket %g mm: forces differ by %g", res, df);
    const int64_t nf = orc_bracket_fixed(nx, ny, nz, nullptr);
    std::vector<int64_t> fx(nf);
    orc_bracket_fixed(nx, ny, nz, fx.data());
    CHECK(fx == lc.getConstrainedDofs(), "bracket %g mm: fixed DOFs", res);
    const std::vector<double> rho0 = dom.initialDensity(0.3, 0.0);
    CHECK((int64_t)rho0.size() == ne, "bracket %g mm: density size", res);
  }
  // a short oracle run (assembly, penalty, PCG, sensitivities, filter, OC) on a small
  // cantilever, once in each density index mode
  const int nx = 8, ny = 4, nz = 3;
  const int64_t nn = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  std::vector<double> F(3 * nn, 0.0);
  F[3 * am::nodeId(nx, ny / 2, 0, ny, nz) + 2] = -1.0;
  std::vector<int64_t> fixed;
  for (int j = 0; j <= ny; ++j)
    for (int k = 0; k <= nz; ++k)
      for (int c = 0; c < 3; ++c) fixed.push_back(3 * am::nodeId(0, j, k, ny, nz) + c);
  for (int ref = 0; ref <= 1; ++ref) {
    orc_config cfg;
    std::memset(&cfg, 0, sizeof(cfg));
    cfg.nx = nx;
    cfg.ny = ny;
    cfg.nz = nz;
    cfg.h = 1.0;
    cfg.E = 1.0;
    cfg.nu = 0.3;
    cfg.volume_fraction = 0.3;
    cfg.penalty = 3.0;
    cfg.filter_radius = 1.5;
    cfg.move_limit = 0.2;
    cfg.rho_min = 0.01;
    cfg.E_min = 1e-9;
    cfg.max_iterations = 4;
    cfg.pcg_tol = 1e-8;
    cfg.pcg_maxit = 10000;
    cfg.penalty_factor = 1e8;
    cfg.ref_compat = ref;
    std::vector<double> rho(nx * ny * nz, 0.3), u(3 * nn);
    std::vector<orc_iter_result> res(4);
    int conv = 0;
    const int it = orc_run(&cfg, rho.data(), F.data(), fixed.data(), (int64_t)fixed.size(), 4,
                           res.data(), &conv, u.data());
    CHECK(it == 4, "oracle run: %d iterations", it);
    for (int t = 0; t < it; ++t)
      CHECK(res[t].solver_converged && std::isfinite(res[t].compliance) &&
                res[t].true_rel_residual <= 1.5e-8,
            "oracle run (ref %d) iteration %d", ref, t);
  }
  // the filter at a radius wider than the grid and the OC update's NaN branch
  {
    const int64_t ne = (int64_t)nx * ny * nz;
    std::vector<int64_t> offs(ne + 1);
    const int64_t tot = orc_build_filter(nx, ny, nz, 5.5, offs.data(), nullptr, nullptr);
    std::vector<int64_t> idx(tot);
    std::vector<double> w(tot), dc(ne, -1.0), rho(ne, 0.5), dcf(ne);
    orc_build_filter(nx, ny, nz, 5.5, offs.data(), idx.data(), w.data());
    orc_filter(ne, offs.data(), idx.data(), w.data(), dc.data(), rho.data(), dcf.data());
    for (int64_t e = 0; e < ne; ++e) CHECK(std::fabs(dcf[e] + 1.0) <= 1e-12, "filter e=%ld", (long)e);
    std::vector<uint8_t> des(ne, 1);
    dc[3] = 0.5;
    int np = 0;
    double lm = 0.0;
    orc_oc_update(ne, rho.data(), dc.data(), des.data(), ne, 0.4, 0.2, 0.01, &np, &lm);
    CHECK(np > 0 && np <= 100 && rho[3] == 0.7, "OC update: %d passes, rho[3] = %g", np, rho[3]);
  }
  std::printf("sanitize_check: %s\n", fails ? "FAILED" : "ok");
  return fails ? 1 : 0;
}
