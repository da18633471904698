// simp_py.hip — the suite's Python SIMP loop (SIMPOptimizer, src/am/numerical/topopt.py:84-411)
// on the GPU, sharing the matrix-free hex8 operator tile with the C++-engine path.
//
// What differs from TopologyOptimizer (amt.hip), phase by phase:
//   filter    cone kernel on the (2 ceil(r)+1)^3 box, normalised to sum 1 (topopt.py:134-151),
//             applied as scipy.ndimage.convolve(mode='reflect') (:153-156): half-sample
//             symmetric extension d c b a | a b c d | d c b a, periodic beyond one reflection;
//             densities are clipped to [rho_min, 1] after it (:302-303), sensitivities are
//             filtered as convolve(dc * rho) / max(rho, 1e-3) (:363)
//   stiffness void elements carry nothing, fixed ones E, design ones E_min + rho~^p (E - E_min)
//             (fem.py:181-207); density layout (nx, ny, nz) C-order, like the scales
//   solve     per load case with its constrained DOFs and the orphan DOFs (|K_ii| < 1e-15)
//             eliminated (fem.py:262-304): the operator is M K M (M = free-DOF mask), so
//             FP64 Jacobi-PCG on it is CG on K_ff; the reference solves K_ff directly, here
//             the solve runs to a relative residual (solve_tol) and reports the FP64 true one
//   compliance sum over non-void elements of E_ratio u_e^T Ke0 u_e (fem.py:356-384), summed
//             over load cases, as are dc = -p rho~^(p-1) u_e^T Ke0 u_e on design elements
//             (topopt.py:158-198, 329-358)
//   OC        bisection from [1e-9, 1e9], break test (l2-l1)/max(l1+l2, 1e-10) <= 1e-3 before
//             each of at most 100 passes, Be = sqrt(-dcf / (lmid + 1e-10)); the density of
//             the last evaluated lmid is kept (topopt.py:200-262)
//
// Every reduction is a fixed-order block partial reduced by the next kernel (device.hpp).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/amt.h"
#include "device.hpp"

namespace amt {
int set_global_error(int code, const std::string& msg);
void hex8_stiffness(double E, double nu, double h, double* Ke);
}  // namespace amt

using namespace amt;

namespace {

constexpr int kSpRedBlocks = 1024;  // grid of the grid-stride kernels (partials per value)
constexpr int kSpChunk = 32;        // PCG iterations between host polls
constexpr int kOcMaxPasses = 100;   // topopt.py:235

struct SpState {  // PCG on K_ff
  double rz, rr, ff, thresh, beta;
  int it, done, converged, maxit;
};

struct SpOc {  // OC bisection
  double l1, l2, lmid, vf, n_design;
  int passes, done;
};

enum Part { P_OP = 0, P_UPD, P_INIT, P_OC, P_FIN, P_COMP, P_COUNT };

// scipy.ndimage 'reflect' index: half-sample symmetric, period 2n.
__device__ __forceinline__ int reflect_index(int m, int n) {
  const int p = 2 * n;
  m %= p;
  if (m < 0) m += p;
  return m < n ? m : p - 1 - m;
}

enum FilterMode { F_PLAIN = 0, F_DENSITY = 1, F_SENS = 2 };

// y[e] = sum_m w[m] x[e + R - m] over the (2R+1)^3 box (convolution; w is symmetric), thread
// per element. F_DENSITY clips to [rho_min, 1]; F_SENS filters x = dc * rho and divides by
// max(rho, 1e-3).
__global__ __launch_bounds__(kBlock) void k_sp_filter(const double* __restrict__ x,
                                                      const double* __restrict__ rho,
                                                      double* __restrict__ y,
                                                      const double* __restrict__ w, int R,
                                                      int nx, int ny, int nz, int mode,
                                                      double rho_min) {
  const int64_t n = (int64_t)nx * ny * nz;
  const int S = 2 * R + 1;
  AMT_GRID_STRIDE(e, n) {
    const int k = (int)(e % nz), j = (int)((e / nz) % ny), i = (int)(e / ((int64_t)nz * ny));
    double acc = 0.0;
    for (int a = 0; a < S; ++a) {
      const int ii = reflect_index(i + R - a, nx);
      for (int b = 0; b < S; ++b) {
        const int jj = reflect_index(j + R - b, ny);
        const double* wr = w + (a * S + b) * S;
        const int64_t row = ((int64_t)ii * ny + jj) * nz;
        for (int c = 0; c < S; ++c) {
          const double wv = wr[c];
          if (wv == 0.0) continue;  // the box corners outside the cone add exact zeros
          const int64_t q = row + reflect_index(k + R - c, nz);
          const double xv = mode == F_SENS ? x[q] * rho[q] : x[q];
          acc = fma(wv, xv, acc);
        }
      }
    }
    if (mode == F_DENSITY) acc = fmin(fmax(acc, rho_min), 1.0);
    if (mode == F_SENS) acc = acc / fmax(rho[e], 1e-3);
    y[e] = acc;
  }
}

// s_e = E_eff / E (fem.py:193-200).
__global__ __launch_bounds__(kBlock) void k_sp_scales(const int8_t* __restrict__ grid,
                                                      const double* __restrict__ rhof,
                                                      double* __restrict__ s, int64_t n,
                                                      double E, double E_min, double pen) {
  AMT_GRID_STRIDE(e, n) {
    const int cell = grid[e];
    s[e] = cell == -1 ? 0.0 : (cell == 1 ? 1.0 : (E_min + pow(rhof[e], pen) * (E - E_min)) / E);
  }
}

// diag(K): thread per DOF, from the eight elements around its node.
__global__ __launch_bounds__(kBlock) void k_sp_diag(const double* __restrict__ s,
                                                    const double* __restrict__ Ke, GridDims g,
                                                    double* __restrict__ diag) {
  AMT_GRID_STRIDE(q, 3 * g.n_nodes) {
    const int64_t n = q / 3;
    const int c = (int)(q % 3);
    int i, j, k;
    node_decode(n, g, i, j, k);
    double s8[8];
    load_element_scales(s, g, i, j, k, s8);
    double d = 0.0;
#pragma unroll
    for (int a = 0; a < 8; ++a) d = fma(s8[a], Ke[(3 * a + c) * 24 + 3 * a + c], d);
    diag[q] = d;
  }
}

// Free-DOF mask of a load case (not constrained, not orphaned) and the Jacobi inverse on it.
__global__ __launch_bounds__(kBlock) void k_sp_free(const double* __restrict__ diag,
                                                    const uint8_t* __restrict__ cons,
                                                    uint8_t* __restrict__ freem,
                                                    double* __restrict__ dinv, int64_t n) {
  AMT_GRID_STRIDE(q, n) {
    const bool f = !cons[q] && !(fabs(diag[q]) < 1e-15);
    freem[q] = f;
    dinv[q] = f ? 1.0 / diag[q] : 0.0;
  }
}

// y = M K M x on the LDS tile (FP64); optional partials of x.y.
__global__ __launch_bounds__(kBlock) void k_sp_op(const double* __restrict__ x,
                                                  double* __restrict__ y,
                                                  const uint8_t* __restrict__ freem,
                                                  const double* __restrict__ s,
                                                  const double* __restrict__ Ke, GridDims g,
                                                  TileGrid tg, const SpState* st,
                                                  double* partials) {
  if (st && st->done) return;
  __shared__ TileSmem<double> sm;
  int i0, j0, k0;
  const bool has = tile_origin(tg, i0, j0, k0);
  if (has) tile_load(sm, g, i0, j0, k0, [&](int64_t q) { return freem[q] ? x[q] : 0.0; }, s);
  __syncthreads();
  int li, lj, lk;
  tile_thread(li, lj, lk);
  const int i = i0 + li, j = j0 + lj, k = k0 + lk;
  double acc[1] = {0.0};
  if (has && i <= g.nx && j <= g.ny && k <= g.nz) {
    const int64_t n = ((int64_t)i * g.nny + j) * g.nnz + k;
    double yv[3], own[3];
    tile_apply(sm, li, lj, lk, Ke, yv, own);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t q = 3 * n + c;
      const double v = freem[q] ? yv[c] : 0.0;
      y[q] = v;
      acc[0] = fma(own[c], v, acc[0]);
    }
  }
  if (partials) store_partials<1>(acc, partials, gridDim.x);
}

// r = M (F - K u), z = D^-1 r, p = z; partials r.z, r.r, (MF).(MF).
__global__ __launch_bounds__(kBlock) void k_sp_init(const double* __restrict__ F,
                                                    const double* __restrict__ Ku,
                                                    const uint8_t* __restrict__ freem,
                                                    const double* __restrict__ dinv,
                                                    double* __restrict__ r,
                                                    double* __restrict__ z,
                                                    double* __restrict__ p, int64_t n,
                                                    double* partials, int stride) {
  double acc[3] = {0.0, 0.0, 0.0};
  AMT_GRID_STRIDE(q, n) {
    const double f = freem[q] ? F[q] : 0.0;
    const double rv = freem[q] ? F[q] - Ku[q] : 0.0;
    const double zv = dinv[q] * rv;
    r[q] = rv;
    z[q] = zv;
    p[q] = zv;
    acc[0] = fma(rv, zv, acc[0]);
    acc[1] = fma(rv, rv, acc[1]);
    acc[2] = fma(f, f, acc[2]);
  }
  store_partials<3>(acc, partials, stride);
}

__global__ void k_sp_start(SpState* st, const double* partials, int nb, int stride, double tol,
                           int maxit) {
  double v[3];
  reduce_partials<3>(partials, nb, stride, v);
  if (threadIdx.x) return;
  SpState s;
  s.rz = v[0];
  s.rr = v[1];
  s.ff = v[2];
  s.thresh = fmax(tol * tol * v[2], 2.2250738585072014e-308);
  s.beta = 0.0;
  s.it = 0;
  s.maxit = maxit;
  s.converged = (v[2] == 0.0) || (v[1] < s.thresh);
  s.done = s.converged || maxit <= 0;
  *st = s;
}

// alpha = rz / p.q; u += alpha p; r -= alpha q; z = D^-1 r; partials r.z, r.r.
__global__ __launch_bounds__(kBlock) void k_sp_update(double* __restrict__ u,
                                                      double* __restrict__ r,
                                                      double* __restrict__ z,
                                                      const double* __restrict__ p,
                                                      const double* __restrict__ q,
                                                      const double* __restrict__ dinv,
                                                      int64_t n, const SpState* st,
                                                      const double* opart, int nbo,
                                                      double* upart, int stride) {
  if (st->done) return;
  double pq[1];
  reduce_partials<1>(opart, nbo, nbo, pq);
  const double a = st->rz / pq[0];
  double acc[2] = {0.0, 0.0};
  AMT_GRID_STRIDE(i, n) {
    u[i] = fma(a, p[i], u[i]);
    const double rv = fma(-a, q[i], r[i]);
    const double zv = dinv[i] * rv;
    r[i] = rv;
    z[i] = zv;
    acc[0] = fma(rv, zv, acc[0]);
    acc[1] = fma(rv, rv, acc[1]);
  }
  store_partials<2>(acc, upart, stride);
}

// One block: convergence test on |r|^2 and beta = rz_new / rz.
__global__ void k_sp_scalar(SpState* st, const double* upart, int nb, int stride) {
  if (st->done) return;
  double v[2];
  reduce_partials<2>(upart, nb, stride, v);
  if (threadIdx.x) return;
  SpState s = *st;
  s.it += 1;
  s.rr = v[1];
  if (v[1] < s.thresh) {
    s.done = 1;
    s.converged = 1;
  } else {
    s.beta = v[0] / s.rz;
    s.rz = v[0];
    if (s.it >= s.maxit) s.done = 1;
  }
  *st = s;
}

__global__ __launch_bounds__(kBlock) void k_sp_pupdate(double* __restrict__ p,
                                                       const double* __restrict__ z, int64_t n,
                                                       const SpState* st) {
  if (st->done) return;
  const double b = st->beta;
  AMT_GRID_STRIDE(i, n) p[i] = fma(b, p[i], z[i]);
}

// Partials of |M (F - K u)|^2 (true residual).
__global__ __launch_bounds__(kBlock) void k_sp_resnorm(const double* __restrict__ F,
                                                       const double* __restrict__ Ku,
                                                       const uint8_t* __restrict__ freem,
                                                       int64_t n, double* partials,
                                                       int stride) {
  double acc[1] = {0.0};
  AMT_GRID_STRIDE(q, n) {
    const double rv = freem[q] ? F[q] - Ku[q] : 0.0;
    acc[0] = fma(rv, rv, acc[0]);
  }
  store_partials<1>(acc, partials, stride);
}

// Thread per element (C order): ce = u_e^T Ke0 u_e; compliance partial of E_ratio ce over
// non-void elements (fem.py:356-384); dc (+)= -p rho~^(p-1) ce on design elements
// (topopt.py:158-198). first: the first load case writes dc, later ones add to it.
__global__ __launch_bounds__(kBlock) void k_sp_sens(const double* __restrict__ u,
                                                    const int8_t* __restrict__ grid,
                                                    const double* __restrict__ rhof,
                                                    const double* __restrict__ Ke, GridDims g,
                                                    double E, double E_min, double pen,
                                                    double* __restrict__ dc, int first,
                                                    double* partials, int stride) {
  double acc[1] = {0.0};
  AMT_GRID_STRIDE(e, g.n_elem) {
    const int k = (int)(e % g.nz), j = (int)((e / g.nz) % g.ny),
              i = (int)(e / ((int64_t)g.nz * g.ny));
    double ue[24];
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int64_t node = ((int64_t)(i + kOX[a]) * g.nny + (j + kOY[a])) * g.nnz + (k + kOZ[a]);
#pragma unroll
      for (int c = 0; c < 3; ++c) ue[3 * a + c] = u[3 * node + c];
    }
    double ce = 0.0;
#pragma unroll
    for (int c = 0; c < 24; ++c) {
      double t = 0.0;
#pragma unroll
      for (int r = 0; r < 24; ++r) t = fma(ue[r], Ke[r * 24 + c], t);
      ce = fma(t, ue[c], ce);
    }
    const int cell = grid[e];
    double d = 0.0;
    if (cell != -1) {
      const double rho = rhof[e];
      const double ratio = cell == 0 ? E_min / E + pow(rho, pen) * (1.0 - E_min / E) : 1.0;
      acc[0] = fma(ratio, ce, acc[0]);
      if (cell == 0) d = -pen * pow(rho, pen - 1.0) * ce;
    }
    dc[e] = first ? d : dc[e] + d;
  }
  store_partials<1>(acc, partials, stride);
}

__device__ __forceinline__ double sp_oc_density(double rho, double dcf, double lmid,
                                                double move, double rho_min) {
  const double Be = sqrt(-dcf / (lmid + 1e-10));
  return fmax(rho_min, fmax(rho - move, fmin(1.0, fmin(rho + move, rho * Be))));
}

__device__ __forceinline__ bool sp_oc_stop(const SpOc& s) {
  return s.passes >= kOcMaxPasses || s.l1 + s.l2 < 1e-12 ||
         (s.l2 - s.l1) / fmax(s.l1 + s.l2, 1e-10) <= 1e-3;
}

// Volume of the design elements at lmid = (l1 + l2) / 2.
__global__ __launch_bounds__(kBlock) void k_sp_oc_pass(const double* __restrict__ rho,
                                                       const double* __restrict__ dcf,
                                                       const int8_t* __restrict__ grid,
                                                       int64_t n, double move, double rho_min,
                                                       const SpOc* st, double* partials,
                                                       int stride) {
  if (st->done) return;
  const double lmid = 0.5 * (st->l1 + st->l2);
  double acc[1] = {0.0};
  AMT_GRID_STRIDE(e, n) {
    if (grid[e] == 0) acc[0] += sp_oc_density(rho[e], dcf[e], lmid, move, rho_min);
  }
  store_partials<1>(acc, partials, stride);
}

__global__ void k_sp_oc_decide(SpOc* st, const double* partials, int nb, int stride) {
  if (st->done) return;
  double v[1];
  reduce_partials<1>(partials, nb, stride, v);
  if (threadIdx.x) return;
  SpOc s = *st;
  s.lmid = 0.5 * (s.l1 + s.l2);
  if (v[0] / s.n_design > s.vf)
    s.l1 = s.lmid;
  else
    s.l2 = s.lmid;
  s.passes += 1;
  s.done = sp_oc_stop(s);
  *st = s;
}

// The density at the last evaluated lmid (non-design elements keep theirs); partials of
// max |rho_new - rho| over all elements and of the design elements' sum.
__global__ __launch_bounds__(kBlock) void k_sp_oc_final(const double* __restrict__ rho,
                                                        const double* __restrict__ dcf,
                                                        const int8_t* __restrict__ grid,
                                                        double* __restrict__ rho_new,
                                                        int64_t n, double move,
                                                        double rho_min, const SpOc* st,
                                                        double* cpart, double* vpart,
                                                        int stride) {
  const double lmid = st->lmid;
  double mx[1] = {0.0}, vol[1] = {0.0};
  AMT_GRID_STRIDE(e, n) {
    const double r = rho[e];
    const bool des = grid[e] == 0;
    const double rn = des ? sp_oc_density(r, dcf[e], lmid, move, rho_min) : r;
    rho_new[e] = rn;
    mx[0] = fmax(mx[0], fabs(rn - r));
    if (des) vol[0] += rn;
  }
  store_partials<1, true>(mx, cpart, stride);
  store_partials<1>(vol, vpart, stride);
}

unsigned nblk(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
unsigned nred(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(nblk(n), kSpRedBlocks));
}

}  // namespace

struct amt_simp {
  GridDims g{};
  TileGrid tg{};
  int64_t n_e = 0, n_dof = 0, n_design = 0;
  amt_simp_params prm{};
  double E = 0.0;
  int R = 0;
  hipStream_t stream = nullptr;
  std::vector<void*> allocs;
  double *Ke = nullptr, *w = nullptr;
  int8_t* grid = nullptr;
  double *rho = nullptr, *rho_new = nullptr, *rhof = nullptr, *s = nullptr, *diag = nullptr;
  double *dc = nullptr, *dcf = nullptr, *tmp = nullptr;
  double *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *dinv = nullptr;
  uint8_t* freem = nullptr;
  double* part = nullptr;
  int stride = 0;
  SpState* st = nullptr;
  SpOc* oc = nullptr;
  double* scal = nullptr;  // finalized scalars
  struct LoadCase {
    double* F = nullptr;
    uint8_t* cons = nullptr;
    double* u = nullptr;
  };
  std::vector<LoadCase> lcs;

  double* region(int r) { return part + (size_t)r * 3 * stride; }
  unsigned nb_tile() const { return 8u * (unsigned)tg.per_xcd; }
};

namespace {

int err(int code, const std::string& m) { return amt::set_global_error(code, m); }

template <typename T>
int dalloc(amt_simp* h, T** ptr, size_t count) {
  void* v = nullptr;
  if (hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess)
    return err(AMT_EHIP, "hipMalloc failed");
  h->allocs.push_back(v);
  *ptr = (T*)v;
  return AMT_OK;
}

int hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) return err(AMT_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return AMT_OK;
}

int launch_ok(const char* what) { return hip_ok(hipGetLastError(), what); }

#define SP_CHECK(call)             \
  do {                             \
    int rc_ = (call);              \
    if (rc_ != AMT_OK) return rc_; \
  } while (0)

// topopt.py:134-151: cone weights on the (2 ceil(r) + 1)^3 box, divided by their sum.
std::vector<double> cone_kernel(double r, int& R) {
  const int size = (int)(2 * std::ceil(r) + 1);
  R = size / 2;
  std::vector<double> w((size_t)size * size * size, 0.0);
  double sum = 0.0;
  for (int i = 0; i < size; ++i)
    for (int j = 0; j < size; ++j)
      for (int k = 0; k < size; ++k) {
        const double d = std::sqrt((double)((i - R) * (i - R) + (j - R) * (j - R) + (k - R) * (k - R)));
        if (d <= r) w[((size_t)i * size + j) * size + k] = r - d;
      }
  for (double v : w) sum += v;
  for (double& v : w) v /= sum;
  return w;
}

void filter(amt_simp* h, const double* x, double* y, int mode) {
  hipLaunchKernelGGL(k_sp_filter, dim3(nred(h->n_e)), dim3(kBlock), 0, h->stream, x, h->rho, y,
                     h->w, h->R, h->g.nx, h->g.ny, h->g.nz, mode, h->prm.rho_min);
}

// PCG on K_ff of load case lc from its previous displacement (0 at the first step).
int solve(amt_simp* h, amt_simp::LoadCase& lc, int* iters, double* rel, bool* conv) {
  const int64_t n = h->n_dof;
  const unsigned nbt = h->nb_tile(), nbr = nred(n);
  hipLaunchKernelGGL(k_sp_free, dim3(nbr), dim3(kBlock), 0, h->stream, h->diag, lc.cons,
                     h->freem, h->dinv, n);
  hipLaunchKernelGGL(k_sp_op, dim3(nbt), dim3(kBlock), 0, h->stream, lc.u, h->q, h->freem,
                     h->s, h->Ke, h->g, h->tg, nullptr, nullptr);
  hipLaunchKernelGGL(k_sp_init, dim3(nbr), dim3(kBlock), 0, h->stream, lc.F, h->q, h->freem,
                     h->dinv, h->r, h->z, h->p, n, h->region(P_INIT), h->stride);
  hipLaunchKernelGGL(k_sp_start, dim3(1), dim3(kBlock), 0, h->stream, h->st,
                     h->region(P_INIT), (int)nbr, h->stride, h->prm.solve_tol,
                     h->prm.solve_maxit);
  SP_CHECK(launch_ok("simp solve init"));
  SpState hs{};
  for (;;) {
    SP_CHECK(hip_ok(hipMemcpyAsync(&hs, h->st, sizeof(SpState), hipMemcpyDeviceToHost,
                                   h->stream), "copy"));
    SP_CHECK(hip_ok(hipStreamSynchronize(h->stream), "sync"));
    if (hs.done) break;
    for (int t = 0; t < kSpChunk; ++t) {
      hipLaunchKernelGGL(k_sp_op, dim3(nbt), dim3(kBlock), 0, h->stream, h->p, h->q, h->freem,
                         h->s, h->Ke, h->g, h->tg, h->st, h->region(P_OP));
      hipLaunchKernelGGL(k_sp_update, dim3(nbr), dim3(kBlock), 0, h->stream, lc.u, h->r, h->z,
                         h->p, h->q, h->dinv, n, h->st, h->region(P_OP), (int)nbt,
                         h->region(P_UPD), h->stride);
      hipLaunchKernelGGL(k_sp_scalar, dim3(1), dim3(kBlock), 0, h->stream, h->st,
                         h->region(P_UPD), (int)nbr, h->stride);
      hipLaunchKernelGGL(k_sp_pupdate, dim3(nbr), dim3(kBlock), 0, h->stream, h->p, h->z, n,
                         h->st);
    }
    SP_CHECK(launch_ok("simp pcg"));
  }
  // FP64 true residual |M (F - K u)| / |M F|
  hipLaunchKernelGGL(k_sp_op, dim3(nbt), dim3(kBlock), 0, h->stream, lc.u, h->q, h->freem,
                     h->s, h->Ke, h->g, h->tg, nullptr, nullptr);
  hipLaunchKernelGGL(k_sp_resnorm, dim3(nbr), dim3(kBlock), 0, h->stream, lc.F, h->q, h->freem,
                     n, h->region(P_INIT), h->stride);
  hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kBlock), 0, h->stream, h->region(P_INIT),
                     (int)nbr, h->stride, h->scal);
  double rr = 0.0;
  SP_CHECK(hip_ok(hipMemcpyAsync(&rr, h->scal, sizeof(double), hipMemcpyDeviceToHost, h->stream),
                  "copy"));
  SP_CHECK(hip_ok(hipStreamSynchronize(h->stream), "sync"));
  *iters = hs.it;
  *rel = hs.ff > 0.0 ? std::sqrt(rr / hs.ff) : 0.0;
  *conv = hs.converged != 0;
  return AMT_OK;
}

}  // namespace

extern "C" {

int amt_simp_create(int nx, int ny, int nz, double element_size, double E, double nu,
                    const int8_t* grid, const amt_simp_params* prm, amt_simp** out) {
  if (!grid || !prm || !out) return err(AMT_EINVAL, "null argument");
  *out = nullptr;
  if (nx <= 0 || ny <= 0 || nz <= 0) return err(AMT_EINVAL, "grid dimensions must be positive");
  if (!(element_size > 0.0) || !(E > 0.0) || !(prm->filter_radius > 0.0) ||
      prm->filter_radius > 16.0)
    return err(AMT_EINVAL, "element size, E and filter radius (<= 16) must be positive");
  const int64_t n_e = (int64_t)nx * ny * nz;
  int64_t n_design = 0;
  for (int64_t e = 0; e < n_e; ++e) {
    if (grid[e] < -1 || grid[e] > 1) return err(AMT_EINVAL, "grid values must be -1, 0 or 1");
    n_design += grid[e] == 0;
  }
  if (n_design == 0) return err(AMT_EINVAL, "no design elements");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
    return err(AMT_EHIP, "no HIP device");
  amt_simp* h = new amt_simp();
  auto bail = [&](int rc) {
    amt_simp_destroy(h);
    return rc;
  };
  h->g.nx = nx;
  h->g.ny = ny;
  h->g.nz = nz;
  h->g.nny = ny + 1;
  h->g.nnz = nz + 1;
  h->g.n_nodes = (int64_t)(nx + 1) * (ny + 1) * (nz + 1);
  h->g.n_elem = n_e;
  h->tg = make_tile_grid(h->g);
  h->n_e = n_e;
  h->n_dof = 3 * h->g.n_nodes;
  h->n_design = n_design;
  h->prm = *prm;
  h->E = E;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(err(AMT_EHIP, "stream creation failed"));
  double Ke[576];
  hex8_stiffness(E, nu, element_size, Ke);
  const std::vector<double> w = cone_kernel(prm->filter_radius, h->R);
  const int64_t nd = h->n_dof;
  h->stride = std::max<int>(kSpRedBlocks, (int)h->nb_tile());
  int rc = AMT_OK;
  if ((rc = dalloc(h, &h->Ke, 576)) || (rc = dalloc(h, &h->w, w.size())) ||
      (rc = dalloc(h, &h->grid, n_e)) || (rc = dalloc(h, &h->rho, n_e)) ||
      (rc = dalloc(h, &h->rho_new, n_e)) || (rc = dalloc(h, &h->rhof, n_e)) ||
      (rc = dalloc(h, &h->s, n_e)) || (rc = dalloc(h, &h->dc, n_e)) ||
      (rc = dalloc(h, &h->dcf, n_e)) || (rc = dalloc(h, &h->tmp, n_e)) ||
      (rc = dalloc(h, &h->diag, nd)) || (rc = dalloc(h, &h->r, nd)) ||
      (rc = dalloc(h, &h->z, nd)) || (rc = dalloc(h, &h->p, nd)) || (rc = dalloc(h, &h->q, nd)) ||
      (rc = dalloc(h, &h->dinv, nd)) || (rc = dalloc(h, &h->freem, nd)) ||
      (rc = dalloc(h, &h->part, (size_t)P_COUNT * 3 * h->stride)) ||
      (rc = dalloc(h, &h->st, 1)) || (rc = dalloc(h, &h->oc, 1)) || (rc = dalloc(h, &h->scal, 8)))
    return bail(rc);
  std::vector<double> rho0(n_e, prm->volume_fraction);  // topopt.py:124
  if (hipMemcpy(h->Ke, Ke, sizeof(Ke), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->w, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->grid, grid, n_e, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(h->rho, rho0.data(), n_e * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return bail(err(AMT_EHIP, "upload failed"));
  *out = h;
  return AMT_OK;
}

int amt_simp_destroy(amt_simp* h) {
  if (!h) return AMT_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (void* v : h->allocs) (void)hipFree(v);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return AMT_OK;
}

int amt_simp_add_load_case(amt_simp* h, const double* F, int64_t n_dof,
                           const int64_t* constrained, int64_t n_constrained) {
  if (!h || !F || (n_constrained > 0 && !constrained) || n_constrained < 0)
    return err(AMT_EINVAL, "null argument");
  if (n_dof != h->n_dof) return err(AMT_EINVAL, "force vector size mismatch");
  std::vector<uint8_t> cons(h->n_dof, 0);
  for (int64_t t = 0; t < n_constrained; ++t) {
    if (constrained[t] < 0 || constrained[t] >= h->n_dof)
      return err(AMT_ERANGE, "constrained DOF out of range");
    cons[constrained[t]] = 1;
  }
  for (int64_t t = 0; t < n_dof; ++t)
    if (!std::isfinite(F[t])) return err(AMT_ENAN, "non-finite force");
  amt_simp::LoadCase lc;
  if (dalloc(h, &lc.F, h->n_dof) || dalloc(h, &lc.cons, h->n_dof) || dalloc(h, &lc.u, h->n_dof))
    return AMT_EHIP;
  if (hipMemcpy(lc.F, F, n_dof * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(lc.cons, cons.data(), n_dof, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(lc.u, 0, n_dof * sizeof(double)) != hipSuccess)
    return err(AMT_EHIP, "upload failed");
  h->lcs.push_back(lc);
  return AMT_OK;
}

int amt_simp_set_density(amt_simp* h, const double* rho, int64_t n_e) {
  if (!h || !rho) return err(AMT_EINVAL, "null argument");
  if (n_e != h->n_e) return err(AMT_EINVAL, "density size mismatch");
  for (int64_t e = 0; e < n_e; ++e)
    if (!std::isfinite(rho[e])) return err(AMT_ENAN, "non-finite density");
  if (hipStreamSynchronize(h->stream) != hipSuccess ||
      hipMemcpy(h->rho, rho, n_e * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return err(AMT_EHIP, "upload failed");
  return AMT_OK;
}

int amt_simp_get_density(amt_simp* h, double* rho, int64_t n_e) {
  if (!h || !rho) return err(AMT_EINVAL, "null argument");
  if (n_e != h->n_e) return err(AMT_EINVAL, "density size mismatch");
  if (hipStreamSynchronize(h->stream) != hipSuccess ||
      hipMemcpy(rho, h->rho, n_e * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return err(AMT_EHIP, "copy failed");
  return AMT_OK;
}

int amt_simp_get_displacement(amt_simp* h, int lc, double* u, int64_t n_dof) {
  if (!h || !u) return err(AMT_EINVAL, "null argument");
  if (lc < 0 || lc >= (int)h->lcs.size()) return err(AMT_ERANGE, "load case out of range");
  if (n_dof != h->n_dof) return err(AMT_EINVAL, "displacement size mismatch");
  if (hipStreamSynchronize(h->stream) != hipSuccess ||
      hipMemcpy(u, h->lcs[lc].u, n_dof * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return err(AMT_EHIP, "copy failed");
  return AMT_OK;
}

int amt_simp_filter(amt_simp* h, const double* x, double* y, int64_t n_e) {
  if (!h || !x || !y) return err(AMT_EINVAL, "null argument");
  if (n_e != h->n_e) return err(AMT_EINVAL, "size mismatch");
  SP_CHECK(hip_ok(hipMemcpyAsync(h->tmp, x, n_e * sizeof(double), hipMemcpyHostToDevice,
                                 h->stream), "upload"));
  filter(h, h->tmp, h->dcf, F_PLAIN);
  SP_CHECK(launch_ok("simp filter"));
  SP_CHECK(hip_ok(hipMemcpyAsync(y, h->dcf, n_e * sizeof(double), hipMemcpyDeviceToHost,
                                 h->stream), "copy"));
  return hip_ok(hipStreamSynchronize(h->stream), "sync");
}

int amt_simp_step(amt_simp* h, amt_simp_iteration* out) {
  if (!h || !out) return err(AMT_EINVAL, "null argument");
  if (h->lcs.empty()) return err(AMT_EINVAL, "no load case");
  const amt_simp_params& P = h->prm;
  const int64_t ne = h->n_e;
  const unsigned nbe = nred(ne);
  // rho~ = clip(filter(rho)); element scales; diag(K)
  filter(h, h->rho, h->rhof, F_DENSITY);
  hipLaunchKernelGGL(k_sp_scales, dim3(nbe), dim3(kBlock), 0, h->stream, h->grid, h->rhof, h->s,
                     ne, h->E, P.E_min, P.penalty);
  hipLaunchKernelGGL(k_sp_diag, dim3(nred(h->n_dof)), dim3(kBlock), 0, h->stream, h->s, h->Ke,
                     h->g, h->diag);
  SP_CHECK(launch_ok("simp stiffness"));
  amt_simp_iteration res{};
  res.solver_converged = 1;
  double comp = 0.0;
  for (size_t l = 0; l < h->lcs.size(); ++l) {
    int it = 0;
    double rel = 0.0;
    bool conv = false;
    SP_CHECK(solve(h, h->lcs[l], &it, &rel, &conv));
    res.solver_iterations += it;
    res.max_true_rel_residual = std::max(res.max_true_rel_residual, rel);
    res.solver_converged &= conv ? 1 : 0;
    hipLaunchKernelGGL(k_sp_sens, dim3(nbe), dim3(kBlock), 0, h->stream, h->lcs[l].u, h->grid,
                       h->rhof, h->Ke, h->g, h->E, P.E_min, P.penalty, h->dc, l == 0 ? 1 : 0,
                       h->region(P_COMP), h->stride);
    hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kBlock), 0, h->stream, h->region(P_COMP),
                       (int)nbe, h->stride, h->scal + 1);
    double c = 0.0;
    SP_CHECK(hip_ok(hipMemcpyAsync(&c, h->scal + 1, sizeof(double), hipMemcpyDeviceToHost,
                                   h->stream), "copy"));
    SP_CHECK(hip_ok(hipStreamSynchronize(h->stream), "sync"));
    comp += c;  // topopt.py:340-342: summed over load cases in order
  }
  // dcf = filter(dc * rho) / max(rho, 1e-3)
  filter(h, h->dc, h->dcf, F_SENS);
  // OC bisection, decided on the device pass by pass
  SpOc init{};
  init.l1 = 1e-9;
  init.l2 = 1e9;
  init.vf = P.volume_fraction;
  init.n_design = (double)h->n_design;
  init.passes = 0;
  init.done = 0;
  init.lmid = 0.5 * (init.l1 + init.l2);
  SP_CHECK(hip_ok(hipMemcpyAsync(h->oc, &init, sizeof(SpOc), hipMemcpyHostToDevice, h->stream),
                  "upload"));
  for (int t = 0; t < kOcMaxPasses; ++t) {
    hipLaunchKernelGGL(k_sp_oc_pass, dim3(nbe), dim3(kBlock), 0, h->stream, h->rho, h->dcf,
                       h->grid, ne, P.move_limit, P.rho_min, h->oc, h->region(P_OC), h->stride);
    hipLaunchKernelGGL(k_sp_oc_decide, dim3(1), dim3(kBlock), 0, h->stream, h->oc,
                       h->region(P_OC), (int)nbe, h->stride);
  }
  hipLaunchKernelGGL(k_sp_oc_final, dim3(nbe), dim3(kBlock), 0, h->stream, h->rho, h->dcf,
                     h->grid, h->rho_new, ne, P.move_limit, P.rho_min, h->oc, h->region(P_FIN),
                     h->region(P_FIN) + h->stride, h->stride);
  hipLaunchKernelGGL((k_finalize<1, true>), dim3(1), dim3(kBlock), 0, h->stream,
                     h->region(P_FIN), (int)nbe, h->stride, h->scal + 2);
  hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kBlock), 0, h->stream,
                     h->region(P_FIN) + h->stride, (int)nbe, h->stride, h->scal + 3);
  SP_CHECK(launch_ok("simp oc"));
  double sc[2];
  SP_CHECK(hip_ok(hipMemcpyAsync(sc, h->scal + 2, sizeof(sc), hipMemcpyDeviceToHost, h->stream),
                  "copy"));
  SP_CHECK(hip_ok(hipStreamSynchronize(h->stream), "sync"));
  std::swap(h->rho, h->rho_new);
  res.compliance = comp;
  res.change = sc[0];
  res.volume = sc[1] / (double)h->n_design;
  *out = res;
  return AMT_OK;
}

}  // extern "C"
