/*
 * oracle.cpp — CPU FP64 restatement of the reference SIMP path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libamt.so and bench.py's CPU baseline.
 * Never linked into, loaded by or called from the product. Every function cites the
 * reference lines (cpp_engine/...) it restates. Built with -ffp-contract=off.
 *
 * Stated deviations from the reference (it cannot be built here: Eigen is absent):
 *  - the linear solver is Jacobi-PCG, not Eigen CG + IncompleteCholesky
 *    (linear_solver.cpp:133-140); the stopping rule is Eigen's conjugate_gradient one;
 *  - dense 24x24 products are plain loops in place of Eigen expressions.
 */
#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// element_kernel.cpp:26-50
const double kGP = 0.5773502691896257;
const double kGauss[8][3] = {{-kGP, -kGP, -kGP}, {+kGP, -kGP, -kGP}, {+kGP, +kGP, -kGP},
                             {-kGP, +kGP, -kGP}, {-kGP, -kGP, +kGP}, {+kGP, -kGP, +kGP},
                             {+kGP, +kGP, +kGP}, {-kGP, +kGP, +kGP}};
const double kNode[8][3] = {{-1, -1, -1}, {+1, -1, -1}, {+1, +1, -1}, {-1, +1, -1},
                            {-1, -1, +1}, {+1, -1, +1}, {+1, +1, +1}, {-1, +1, +1}};
// element_kernel.cpp:196-199 / global_assembler.cpp:29-32
const int kOff[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                        {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

inline double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// global_assembler.hpp:171-173
inline double simp(double rho, double p, double Emin) {
  return Emin + std::pow(rho, p) * (1.0 - Emin);
}

struct Csr {
  int64_t n = 0;
  std::vector<int64_t> rowptr;
  std::vector<int32_t> col;
  std::vector<double> val;
};

struct Trip {
  int32_t r, c;
  double v;
};

// global_assembler.cpp:61-222: thread-local triplets over elements in C order
// (static schedule over i keeps the merged order = element order), then a stable
// bucket by row + stable sort by column, summing duplicates in insertion order.
void assemble(int nx, int ny, int nz, const double* Ke0, const double* rho, double penal,
              double Emin, double rho_min, int ref_compat, Csr& K) {
  const int64_t n_dof = 3LL * (nx + 1) * (ny + 1) * (nz + 1);
  int nthr = 1;
#ifdef _OPENMP
  nthr = omp_get_max_threads();
#endif
  std::vector<std::vector<Trip>> local(nthr);
#pragma omp parallel
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    std::vector<Trip>& L = local[tid];
    int64_t dofs[24];
#pragma omp for schedule(static)
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          const int64_t ei = ref_compat ? (int64_t)i * ny * nz + (int64_t)j * nz + k
                                        : (int64_t)k * nx * ny + (int64_t)j * nx + i;
          const double r = rho[ei];
          if (r < rho_min) continue;  // global_assembler.cpp:147
          const double s = simp(r, penal, Emin);
          orc_element_dofs(i, j, k, nx, ny, nz, dofs);
          for (int a = 0; a < 24; ++a)
            for (int b = 0; b < 24; ++b) {
              const double v = s * Ke0[a * 24 + b];
              if (std::fabs(v) > 1e-20) L.push_back({(int32_t)dofs[a], (int32_t)dofs[b], v});
            }
        }
  }
  // every row gets an explicit 0.0 diagonal first, so a row that no element touches still
  // has the slot that coeffRef would insert for a penalty (linear_solver.cpp:80); the
  // summed value of a touched diagonal is unchanged (0.0 + v == v).
  std::vector<int64_t> cnt(n_dof + 1, 0);
  for (int64_t r = 0; r < n_dof; ++r) cnt[r + 1] = 1;
  for (auto& L : local)
    for (auto& t : L) cnt[t.r + 1]++;
  for (int64_t r = 0; r < n_dof; ++r) cnt[r + 1] += cnt[r];
  std::vector<Trip> bucket(cnt[n_dof]);
  {
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (int64_t r = 0; r < n_dof; ++r) bucket[pos[r]++] = {(int32_t)r, (int32_t)r, 0.0};
    for (auto& L : local) {
      for (auto& t : L) bucket[pos[t.r]++] = t;
      std::vector<Trip>().swap(L);
    }
  }
  K.n = n_dof;
  K.rowptr.assign(n_dof + 1, 0);
  std::vector<int64_t> rlen(n_dof, 0);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n_dof; ++r) {
    Trip* b = bucket.data() + cnt[r];
    Trip* e = bucket.data() + cnt[r + 1];
    std::stable_sort(b, e, [](const Trip& x, const Trip& y) { return x.c < y.c; });
    int64_t u = 0;
    for (Trip* t = b; t < e; ++t)
      if (t == b || t->c != (t - 1)->c) ++u;
    rlen[r] = u;
  }
  for (int64_t r = 0; r < n_dof; ++r) K.rowptr[r + 1] = K.rowptr[r] + rlen[r];
  K.col.resize(K.rowptr[n_dof]);
  K.val.resize(K.rowptr[n_dof]);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n_dof; ++r) {
    int64_t o = K.rowptr[r] - 1;
    for (int64_t q = cnt[r]; q < cnt[r + 1]; ++q) {
      const Trip& t = bucket[q];
      if (q == cnt[r] || t.c != bucket[q - 1].c) {
        ++o;
        K.col[o] = t.c;
        K.val[o] = t.v;
      } else {
        K.val[o] += t.v;
      }
    }
  }
}

void spmv(const Csr& K, const double* x, double* y) {
  orc_csr_spmv(K.n, K.rowptr.data(), K.col.data(), K.val.data(), x, y);
}

double dot(int64_t n, const double* a, const double* b) {
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// topology_optimizer.hpp:227-232
inline void slot_decode(int64_t e, int nx, int ny, int& ix, int& iy, int& iz) {
  iz = (int)(e / ((int64_t)nx * ny));
  int64_t rem = e % ((int64_t)nx * ny);
  iy = (int)(rem / nx);
  ix = (int)(rem % nx);
}

}  // namespace

extern "C" {

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// element_kernel.cpp:71-101 (C) and :103-179 (B, Ke0)
void orc_ke0(double E, double nu, double h, double* Ke) {
  double C[6][6] = {{0}};
  const double f = E / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double dg = 1.0 - nu, od = nu, sh = (1.0 - 2.0 * nu) / 2.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) C[i][j] = (i == j ? dg : od) * f;
  for (int i = 3; i < 6; ++i) C[i][i] = sh * f;
  const double a = h / 2.0, inv_a = 1.0 / a, detJ = a * a * a;
  std::memset(Ke, 0, sizeof(double) * 576);
  for (int g = 0; g < 8; ++g) {
    const double xi = kGauss[g][0], eta = kGauss[g][1], zeta = kGauss[g][2];
    double B[6][24] = {{0}};
    for (int n = 0; n < 8; ++n) {
      const double xn = kNode[n][0], yn = kNode[n][1], zn = kNode[n][2];
      const double dxi = 0.125 * xn * (1.0 + yn * eta) * (1.0 + zn * zeta);
      const double deta = 0.125 * yn * (1.0 + xn * xi) * (1.0 + zn * zeta);
      const double dzeta = 0.125 * zn * (1.0 + xn * xi) * (1.0 + yn * eta);
      const double dx = inv_a * dxi, dy = inv_a * deta, dz = inv_a * dzeta;
      const int c = 3 * n;
      B[0][c] = dx;
      B[1][c + 1] = dy;
      B[2][c + 2] = dz;
      B[3][c] = dy;
      B[3][c + 1] = dx;
      B[4][c + 1] = dz;
      B[4][c + 2] = dy;
      B[5][c] = dz;
      B[5][c + 2] = dx;
    }
    double BtC[24][6];
    for (int r = 0; r < 24; ++r)
      for (int q = 0; q < 6; ++q) {
        double s = 0.0;
        for (int m = 0; m < 6; ++m) s += B[m][r] * C[m][q];
        BtC[r][q] = s;
      }
    for (int r = 0; r < 24; ++r)
      for (int q = 0; q < 24; ++q) {
        double s = 0.0;
        for (int m = 0; m < 6; ++m) s += BtC[r][m] * B[m][q];
        Ke[r * 24 + q] += s * detJ;
      }
  }
}

// element_kernel.cpp:186-216
void orc_element_dofs(int i, int j, int k, int nx, int ny, int nz, int64_t* dofs) {
  (void)nx;
  const int64_t ny1 = ny + 1, nz1 = nz + 1;
  for (int n = 0; n < 8; ++n) {
    const int64_t node = (int64_t)(i + kOff[n][0]) * ny1 * nz1 + (int64_t)(j + kOff[n][1]) * nz1 +
                         (k + kOff[n][2]);
    dofs[3 * n] = 3 * node;
    dofs[3 * n + 1] = 3 * node + 1;
    dofs[3 * n + 2] = 3 * node + 2;
  }
}

int64_t orc_assemble_csr(int nx, int ny, int nz, double h, double E, double nu, const double* rho,
                         double penal, double Emin, double rho_min, int ref_compat,
                         int64_t* rowptr, int32_t* col, double* val) {
  if (nx <= 0 || ny <= 0 || nz <= 0) return -1;
  double Ke[576];
  orc_ke0(E, nu, h, Ke);  // recomputed per call, as global_assembler.cpp:86-87
  Csr K;
  assemble(nx, ny, nz, Ke, rho, penal, Emin, rho_min, ref_compat, K);
  if (rowptr) std::memcpy(rowptr, K.rowptr.data(), sizeof(int64_t) * (K.n + 1));
  if (col) std::memcpy(col, K.col.data(), sizeof(int32_t) * K.col.size());
  if (val) std::memcpy(val, K.val.data(), sizeof(double) * K.val.size());
  return (int64_t)K.col.size();
}

void orc_csr_spmv(int64_t n, const int64_t* rowptr, const int32_t* col, const double* val,
                  const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < n; ++r) {
    double s = 0.0;
    for (int64_t q = rowptr[r]; q < rowptr[r + 1]; ++q) s += val[q] * x[col[q]];
    y[r] = s;
  }
}

// linear_solver.cpp:17-28 (findMaxDiagonal) and :35-91 (applyBoundaryConditions).
// The pattern always holds the diagonal of an assembled row; a row with no entry at
// all (every adjacent element skipped) gets P only, as coeffRef would insert it.
double orc_apply_penalty(int64_t n, const int64_t* rowptr, const int32_t* col, double* val,
                         double* F, const int64_t* fixed, int64_t nfixed, double pf) {
  if (nfixed == 0) return 0.0;
  double maxd = 0.0;
  for (int64_t r = 0; r < n; ++r)
    for (int64_t q = rowptr[r]; q < rowptr[r + 1]; ++q)
      if (col[q] == r && std::fabs(val[q]) > maxd) maxd = std::fabs(val[q]);
  const double P = maxd * pf;
  for (int64_t t = 0; t < nfixed; ++t)
    if (fixed[t] < 0 || fixed[t] >= n) return -1.0;
  for (int64_t t = 0; t < nfixed; ++t) {
    const int64_t i = fixed[t];
    int64_t q = rowptr[i];
    while (q < rowptr[i + 1] && col[q] != i) ++q;
    if (q < rowptr[i + 1]) val[q] += P;
    F[i] = 0.0 * P;  // prescribed value 0 (linear_solver.cpp:83)
  }
  return P;
}

// Eigen conjugate_gradient loop (as called from linear_solver.cpp:136-155) with a Jacobi
// preconditioner in place of IncompleteCholesky; zero diagonals get z=r (Eigen's
// DiagonalPreconditioner uses 1 there).
static int pcg_impl(const Csr& K, const double* b, double* x, double tol, int maxit, int* iters,
                    double* rel_err, double* true_rel) {
  const int64_t n = K.n;
  std::vector<double> invd(n, 1.0), r(n), p(n), z(n), q(n);
  for (int64_t i = 0; i < n; ++i)
    for (int64_t t = K.rowptr[i]; t < K.rowptr[i + 1]; ++t)
      if (K.col[t] == i && K.val[t] != 0.0) invd[i] = 1.0 / K.val[t];
  std::fill(x, x + n, 0.0);
  const double rhs2 = dot(n, b, b);
  *iters = 0;
  *rel_err = 0.0;
  if (true_rel) *true_rel = 0.0;
  if (rhs2 == 0.0) return 1;
  const double thresh = std::max(tol * tol * rhs2, DBL_MIN);
  for (int64_t i = 0; i < n; ++i) r[i] = b[i];
  double res2 = rhs2;
  int it = 0;
  bool conv = res2 < thresh;
  if (!conv) {
    for (int64_t i = 0; i < n; ++i) p[i] = invd[i] * r[i];
    double absnew = dot(n, r.data(), p.data());
    while (it < maxit) {
      spmv(K, p.data(), q.data());
      const double alpha = absnew / dot(n, p.data(), q.data());
      for (int64_t i = 0; i < n; ++i) {
        x[i] += alpha * p[i];
        r[i] -= alpha * q[i];
      }
      res2 = dot(n, r.data(), r.data());
      if (res2 < thresh) {
        conv = true;
        break;
      }
      for (int64_t i = 0; i < n; ++i) z[i] = invd[i] * r[i];
      const double absold = absnew;
      absnew = dot(n, r.data(), z.data());
      const double beta = absnew / absold;
      for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
      ++it;
    }
  }
  *iters = it;
  *rel_err = std::sqrt(res2 / rhs2);
  if (true_rel) {
    spmv(K, x, q.data());
    double t2 = 0.0;
    for (int64_t i = 0; i < n; ++i) t2 += (b[i] - q[i]) * (b[i] - q[i]);
    *true_rel = std::sqrt(t2 / rhs2);
  }
  return conv ? 1 : 0;
}

int orc_pcg_jacobi(int64_t n, const int64_t* rowptr, const int32_t* col, const double* val,
                   const double* b, double* x, double tol, int maxit, int* iters,
                   double* rel_err, double* true_rel) {
  Csr K;
  K.n = n;
  K.rowptr.assign(rowptr, rowptr + n + 1);
  K.col.assign(col, col + rowptr[n]);
  K.val.assign(val, val + rowptr[n]);
  return pcg_impl(K, b, x, tol, maxit, iters, rel_err, true_rel);
}

void orc_matfree_apply(int nx, int ny, int nz, const double* Ke, const double* s_c,
                       const double* x, double* y) {
  const int64_t n_dof = 3LL * (nx + 1) * (ny + 1) * (nz + 1);
  std::fill(y, y + n_dof, 0.0);
  int64_t dofs[24];
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const double s = s_c[(int64_t)i * ny * nz + (int64_t)j * nz + k];
        if (s == 0.0) continue;
        orc_element_dofs(i, j, k, nx, ny, nz, dofs);
        for (int a = 0; a < 24; ++a) {
          double t = 0.0;
          for (int b = 0; b < 24; ++b) t += Ke[a * 24 + b] * x[dofs[b]];
          y[dofs[a]] += s * t;
        }
      }
}

// topology_optimizer.cpp:126-161
void orc_sensitivities(int nx, int ny, int nz, const double* Ke, const double* u,
                       const double* rho, double penal, double* dc) {
  const int64_t n_e = (int64_t)nx * ny * nz;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n_e; ++e) {
    int ix, iy, iz;
    slot_decode(e, nx, ny, ix, iy, iz);
    int64_t dofs[24];
    orc_element_dofs(ix, iy, iz, nx, ny, nz, dofs);
    double ue[24];
    for (int a = 0; a < 24; ++a) ue[a] = u[dofs[a]];
    double ce = 0.0;
    for (int b = 0; b < 24; ++b) {
      double t = 0.0;
      for (int a = 0; a < 24; ++a) t += ue[a] * Ke[a * 24 + b];
      ce += t * ue[b];
    }
    dc[e] = -penal * std::pow(rho[e], penal - 1) * ce;
  }
}

// topology_optimizer.cpp:60-119 (dz, then dy, then dx; x-fastest neighbour index)
int64_t orc_build_filter(int nx, int ny, int nz, double r, int64_t* offs, int64_t* idx,
                         double* w) {
  const int R = (int)std::ceil(r);
  const int64_t n_e = (int64_t)nx * ny * nz;
  int64_t tot = 0;
  for (int64_t e = 0; e < n_e; ++e) {
    int ix, iy, iz;
    slot_decode(e, nx, ny, ix, iy, iz);
    if (offs) offs[e] = tot;
    for (int dz = -R; dz <= R; ++dz) {
      const int cz = iz + dz;
      if (cz < 0 || cz >= nz) continue;
      for (int dy = -R; dy <= R; ++dy) {
        const int cy = iy + dy;
        if (cy < 0 || cy >= ny) continue;
        for (int dx = -R; dx <= R; ++dx) {
          const int cx = ix + dx;
          if (cx < 0 || cx >= nx) continue;
          const double dist = std::sqrt((double)(dx * dx + dy * dy + dz * dz));
          if (dist <= r) {
            if (idx) idx[tot] = (int64_t)cz * ny * nx + (int64_t)cy * nx + cx;
            if (w) w[tot] = std::max(0.0, r - dist);
            ++tot;
          }
        }
      }
    }
  }
  if (offs) offs[n_e] = tot;
  return tot;
}

// topology_optimizer.cpp:168-198
void orc_filter(int64_t n_e, const int64_t* offs, const int64_t* idx, const double* w,
                const double* dc, const double* rho, double* dcf) {
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n_e; ++e) {
    double num = 0.0, den = 0.0;
    for (int64_t t = offs[e]; t < offs[e + 1]; ++t) {
      const double rn = rho[idx[t]], dn = dc[idx[t]];
      num += w[t] * rn * dn;
      den += w[t] * rn;
    }
    dcf[e] = den > 1e-15 ? num / den : dc[e];
  }
}

// topology_optimizer.cpp:205-287. std::min/std::max keep their NaN behaviour.
double orc_oc_update(int64_t n_e, double* rho, const double* dcf, const uint8_t* design,
                     int64_t n_design, double vf, double move, double rho_min, int* n_pass,
                     double* lmid_out) {
  double l1 = 1e-20, l2 = 1e20, lmid = 0.0;
  std::vector<double> rn(n_e);
  if (n_design == 0) n_design = n_e;
  int pass = 0;
  for (int it = 0; it < 100; ++it) {
    if ((l2 - l1) / (l1 + l2) < 1e-4) break;
    lmid = 0.5 * (l1 + l2);
    double vol = 0.0;
    for (int64_t e = 0; e < n_e; ++e) {
      const double r = rho[e];
      if (!design[e]) {
        rn[e] = r;
        continue;
      }
      const double Be = std::sqrt(-dcf[e] / (lmid * 1.0 + 1e-20));
      double c = r * Be;
      c = std::max(rho_min, std::max(r - move, std::min(1.0, std::min(r + move, c))));
      rn[e] = c;
      vol += c;
    }
    ++pass;
    if (vol / n_design > vf)
      l1 = lmid;
    else
      l2 = lmid;
  }
  double ch = 0.0;
  for (int64_t e = 0; e < n_e; ++e) {
    const double d = std::fabs(rn[e] - rho[e]);
    if (d > ch) ch = d;
  }
  std::memcpy(rho, rn.data(), sizeof(double) * n_e);
  if (n_pass) *n_pass = pass;
  if (lmid_out) *lmid_out = lmid;
  return ch;
}

// topology_optimizer.cpp:391-520 (run) and :294-384 (runIteration)
int orc_run(const orc_config* cfg, double* rho, const double* F, const int64_t* fixed,
            int64_t nfixed, int max_it, orc_iter_result* res, int* converged, double* u_out) {
  const int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
  const int64_t n_e = (int64_t)nx * ny * nz;
  const int64_t n_dof = 3LL * (nx + 1) * (ny + 1) * (nz + 1);
  double Ke[576];
  orc_ke0(cfg->E, cfg->nu, cfg->h, Ke);
  const int64_t nf = orc_build_filter(nx, ny, nz, cfg->filter_radius, nullptr, nullptr, nullptr);
  std::vector<int64_t> offs(n_e + 1), fidx(nf);
  std::vector<double> fw(nf);
  orc_build_filter(nx, ny, nz, cfg->filter_radius, offs.data(), fidx.data(), fw.data());
  // design mask from the initial density (topology_optimizer.cpp:419-428)
  std::vector<uint8_t> design(n_e);
  int64_t n_design = 0;
  for (int64_t e = 0; e < n_e; ++e) {
    design[e] = (rho[e] > cfg->rho_min + 1e-6) && (rho[e] < 0.99);
    n_design += design[e];
  }
  std::vector<double> u(n_dof), Fb(n_dof), dc(n_e), dcf(n_e);
  *converged = 0;
  int done = 0;
  const int cap = std::min(max_it, cfg->max_iterations);
  for (int iter = 0; iter < cap; ++iter) {
    orc_iter_result R;
    std::memset(&R, 0, sizeof(R));
    double t0 = now_s();
    Csr K;
    assemble(nx, ny, nz, Ke, rho, cfg->penalty, cfg->E_min, cfg->rho_min, cfg->ref_compat, K);
    double t1 = now_s();
    std::memcpy(Fb.data(), F, sizeof(double) * n_dof);
    if (orc_apply_penalty(n_dof, K.rowptr.data(), K.col.data(), K.val.data(), Fb.data(), fixed,
                          nfixed, cfg->penalty_factor) < 0)
      return -1;
    double rel = 0.0, trel = 0.0;
    R.solver_converged = pcg_impl(K, Fb.data(), u.data(), cfg->pcg_tol, cfg->pcg_maxit,
                                  &R.solver_iterations, &rel, &trel);
    R.true_rel_residual = trel;
    double t2 = now_s();
    R.compliance = dot(n_dof, u.data(), F);
    orc_sensitivities(nx, ny, nz, Ke, u.data(), rho, cfg->penalty, dc.data());
    double t3 = now_s();
    orc_filter(n_e, offs.data(), fidx.data(), fw.data(), dc.data(), rho, dcf.data());
    R.max_density_change =
        orc_oc_update(n_e, rho, dcf.data(), design.data(), n_design, cfg->volume_fraction,
                      cfg->move_limit, cfg->rho_min, &R.bisect_passes, nullptr);
    double t4 = now_s();
    double vs = 0.0;
    int64_t dcount = 0;
    for (int64_t e = 0; e < n_e; ++e)
      if (design[e]) {
        vs += rho[e];
        ++dcount;
      }
    R.volume_fraction = dcount > 0 ? vs / dcount : vs / n_e;
    R.assembly_time_s = t1 - t0;
    R.solve_time_s = t2 - t1;
    R.sensitivity_time_s = t3 - t2;
    R.update_time_s = t4 - t3;
    res[iter] = R;
    done = iter + 1;
    if (cfg->verbose)
      std::printf("%6d %14.6e %10.6f %12.4e %8d\n", iter, R.compliance, R.volume_fraction,
                  R.max_density_change, R.solver_iterations);
    if (R.max_density_change < cfg->convergence_tol) {
      *converged = 1;
      break;
    }
  }
  if (u_out) std::memcpy(u_out, u.data(), sizeof(double) * n_dof);
  return done;
}

// am_topopt.cpp:98-239 (BracketDomain), default Config geometry (am_topopt.cpp:47-80)
int64_t orc_bracket_domain(double res, int8_t* grid, int* pnx, int* pny, int* pnz) {
  const int nx = (int)(120.0 / res), ny = (int)(60.0 / res), nz = (int)(80.0 / res);
  *pnx = nx;
  *pny = ny;
  *pnz = nz;
  const int64_t n = (int64_t)nx * ny * nz;
  if (!grid) return n;
  auto idx = [&](int i, int j, int k) {
    return (int64_t)k * ny * nx + (int64_t)j * nx + i;
  };
  std::fill(grid, grid + n, (int8_t)0);
  auto cyl = [&](double cx, double cy, int z0, double rad, int hgt, int8_t v) {
    const int i0 = std::max(0, (int)(cx - rad - 1)), i1 = std::min(nx, (int)(cx + rad + 2));
    const int j0 = std::max(0, (int)(cy - rad - 1)), j1 = std::min(ny, (int)(cy + rad + 2));
    for (int i = i0; i < i1; ++i)
      for (int j = j0; j < j1; ++j) {
        const double di = i - cx, dj = j - cy;
        if (di * di + dj * dj <= rad * rad)
          for (int k = z0; k < std::min(nz, z0 + hgt); ++k) {
            int8_t& g = grid[idx(i, j, k)];
            if (v == -1 || g != -1) g = v;
          }
      }
  };
  const int base_z = std::max(1, (int)(5.0 / res));
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < base_z && k < nz; ++k) grid[idx(i, j, k)] = 1;
  const double hole_r = 8.0 / 2.0 / res;
  const int hox = (int)(15.0 / res), hoy = (int)(15.0 / res);
  const int hp[4][2] = {{hox, hoy}, {nx - hox, hoy}, {hox, ny - hoy}, {nx - hox, ny - hoy}};
  for (auto& p : hp) cyl(p[0], p[1], 0, hole_r, base_z, -1);
  const int ezo = (int)(15.0 / res);
  const int eh = std::max(1, (int)(20.0 / res));
  const int ezc = nz - ezo, ex = nx / 2, ey = ny / 2;
  const int z0 = std::max(0, ezc - eh / 2), z1 = std::min(nz, z0 + eh);
  if (z1 - z0 > 0) {
    cyl(ex, ey, z0, (12.0 / 2.0 + 4.0) / res, z1 - z0, 1);
    cyl(ex, ey, z0, 12.0 / 2.0 / res, z1 - z0, -1);
  }
  const double pr = (12.0 / 2.0 + 4.0) / res;
  if (z0 > base_z) cyl(ex, ey, base_z, pr, z0 - base_z, 1);
  return n;
}

// am_topopt.cpp:248-314
int orc_bracket_forces(int nx, int ny, int nz, double res, double load, double* F) {
  const int64_t n_dof = 3LL * (nx + 1) * (ny + 1) * (nz + 1);
  std::fill(F, F + n_dof, 0.0);
  const int ezo = (int)(15.0 / res);
  const int eh = std::max(1, (int)(20.0 / res));
  const int ezc = nz - ezo;
  const int z0 = std::max(0, ezc - eh / 2), z1 = std::min(nz, z0 + eh);
  int ez = (z0 + z1) / 2;
  ez = std::max(1, std::min(nz - 1, ez));
  const int ex = nx / 2, ey = ny / 2;
  const double margin = res * 1.25;
  double inner_mm = 12.0 / 2.0 + margin;
  const double outer_mm = 12.0 / 2.0 + 4.0;
  if (inner_mm >= outer_mm) inner_mm = 12.0 / 2.0 + 4.0 * 0.5;
  const double ri = inner_mm / res, ro = outer_mm / res;
  std::vector<int64_t> nodes;
  const int search = (int)ro + 2;
  for (int i = std::max(0, ex - search); i <= std::min(nx, ex + search); ++i)
    for (int j = std::max(0, ey - search); j <= std::min(ny, ey + search); ++j) {
      const double dx = i - ex, dy = j - ey;
      const double d = std::sqrt(dx * dx + dy * dy);
      if (d > ri && d <= ro)
        nodes.push_back((int64_t)i * (ny + 1) * (nz + 1) + (int64_t)j * (nz + 1) + ez);
    }
  if (nodes.empty())
    nodes.push_back((int64_t)ex * (ny + 1) * (nz + 1) + (int64_t)ey * (nz + 1) + ez);
  const double fpn = load / nodes.size();
  for (auto nd : nodes) F[nd * 3 + 2] = -fpn;
  return (int)nodes.size();
}

// am_topopt.cpp:319-334
int64_t orc_bracket_fixed(int nx, int ny, int nz, int64_t* fixed) {
  int64_t c = 0;
  for (int i = 0; i <= nx; ++i)
    for (int j = 0; j <= ny; ++j) {
      const int64_t nd = (int64_t)i * (ny + 1) * (nz + 1) + (int64_t)j * (nz + 1);
      for (int q = 0; q < 3; ++q) {
        if (fixed) fixed[c] = nd * 3 + q;
        ++c;
      }
    }
  return c;
}

}  // extern "C"
