// kernels.hpp — gfx950 HIP kernels of the single-problem SIMP hot path (device code only;
// the shared building blocks and the data layout notes are in device.hpp).
#pragma once

#include "device.hpp"

namespace amt {

// ---------------------------------------------------------------------------------------
// SIMP element scale (global_assembler.cpp:141-158 + global_assembler.hpp:171-173)
// ---------------------------------------------------------------------------------------
// s[C-order (i,j,k)] = 0 if rho < rho_min else Emin + rho^p (1-Emin), with rho read at
// C-order (reference) or x-fastest (consistent) position of element (i,j,k).
__global__ void k_simp(const double* __restrict__ rho, double* __restrict__ s64,
                       float* __restrict__ s32, GridDims g, int ref_compat, double penal,
                       double Emin, double rho_min) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= g.n_elem) return;
  int64_t src = e;
  if (!ref_compat) {
    const int k = (int)(e % g.nz);
    const int64_t t = e / g.nz;
    const int j = (int)(t % g.ny);
    const int i = (int)(t / g.ny);
    src = (int64_t)k * g.nx * g.ny + (int64_t)j * g.nx + i;
  }
  const double r = rho[src];
  const double s = (r < rho_min) ? 0.0 : Emin + pow(r, penal) * (1.0 - Emin);
  s64[e] = s;
  s32[e] = (float)s;
}


// y = K x (+ penalty rows when P != null): fixcnt[dof] * P * x[dof]
template <typename T, typename TK>
__global__ __launch_bounds__(kBlock) void k_apply(const T* __restrict__ x, T* __restrict__ y,
                                                  const T* __restrict__ s,
                                                  const TK* __restrict__ Ke,
                                                  const uint8_t* __restrict__ fixcnt,
                                                  const double* __restrict__ P, GridDims g,
                                                  TileGrid tg) {
  __shared__ TileSmem<T> sm;
  int i0, j0, k0;
  if (!tile_origin(tg, i0, j0, k0)) return;  // whole block: no barrier follows for it
  tile_load(sm, g, i0, j0, k0, [&](int64_t q) { return x[q]; }, s);
  __syncthreads();
  int li, lj, lk;
  tile_thread(li, lj, lk);
  const int i = i0 + li, j = j0 + lj, k = k0 + lk;
  if (i > g.nx || j > g.ny || k > g.nz) return;
  const int64_t n = ((int64_t)i * g.nny + j) * g.nnz + k;
  Nbr<T> nb;
  T s8[8];
  tile_fetch(sm, li, lj, lk, nb, s8);
  T yv[3];
  apply_node(nb, s8, Ke, yv);
  if (P) {
    const T pen = (T)(*P);
#pragma unroll
    for (int c = 0; c < 3; ++c) yv[c] = fma((T)fixcnt[3 * n + c] * pen, nb.v[13][c], yv[c]);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) y[3 * n + c] = yv[c];
}

// r = b - K x (with penalty); block partials of |r|^2 (one per block of the tile grid).
__global__ __launch_bounds__(kBlock) void k_residual(const double* __restrict__ x,
                                                     const double* __restrict__ b,
                                                     double* __restrict__ r,
                                                     const double* __restrict__ s,
                                                     const double* __restrict__ Ke,
                                                     const uint8_t* __restrict__ fixcnt,
                                                     const double* __restrict__ P, GridDims g,
                                                     TileGrid tg, double* partials) {
  __shared__ TileSmem<double> sm;
  int i0, j0, k0;
  const bool has = tile_origin(tg, i0, j0, k0);
  if (has) tile_load(sm, g, i0, j0, k0, [&](int64_t q) { return x[q]; }, s);
  __syncthreads();
  int li, lj, lk;
  tile_thread(li, lj, lk);
  const int i = i0 + li, j = j0 + lj, k = k0 + lk;
  double acc[1] = {0.0};
  if (has && i <= g.nx && j <= g.ny && k <= g.nz) {
    const int64_t n = ((int64_t)i * g.nny + j) * g.nnz + k;
    Nbr<double> nb;
    double s8[8];
    tile_fetch(sm, li, lj, lk, nb, s8);
    double yv[3];
    apply_node(nb, s8, Ke, yv);
    const double pen = *P;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      yv[c] = fma((double)fixcnt[3 * n + c] * pen, nb.v[13][c], yv[c]);
      const double rv = b[3 * n + c] - yv[c];
      r[3 * n + c] = rv;
      acc[0] = fma(rv, rv, acc[0]);
    }
  }
  store_partials<1>(acc, partials, gridDim.x);
}

// Diagonal of K without penalty: d = sum_a s8[a] Ke0[3a+c,3a+c]; block partials of max|d|.
__global__ __launch_bounds__(kBlock) void k_diag(double* __restrict__ diag,
                                                 const double* __restrict__ s,
                                                 const double* __restrict__ Ke, GridDims g,
                                                 double* partials, int stride) {
  const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double m[1] = {0.0};
  if (n < g.n_nodes) {
    int i, j, k;
    node_decode(n, g, i, j, k);
    double s8[8];
    load_element_scales(s, g, i, j, k, s8);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double d = 0.0;
#pragma unroll
      for (int a = 0; a < 8; ++a) d = fma(s8[a], Ke[(3 * a + c) * 24 + 3 * a + c], d);
      diag[3 * n + c] = d;
      m[0] = fmax(m[0], fabs(d));
    }
  }
  store_partials<1, true>(m, partials, stride);
}

// P = pf * max|K_ii| (linear_solver.cpp:46-47); Jacobi inverse of the penalised diagonal.
// A zero diagonal gets 1, as Eigen's DiagonalPreconditioner does; such rows carry r = 0.
__global__ void k_dinv(const double* __restrict__ diag, const uint8_t* __restrict__ fixcnt,
                       const double* __restrict__ maxdiag, double pf, double* __restrict__ P,
                       double* __restrict__ dinv, float* __restrict__ dinv32, int64_t n_dof) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const double pen = pf * maxdiag[0];
  if (i == 0) P[0] = pen;
  if (i >= n_dof) return;
  double d = diag[i];
  for (int t = 0; t < fixcnt[i]; ++t) d += pen;  // one K_ii += P per listed occurrence
  const double v = (d != 0.0) ? 1.0 / d : 1.0;
  dinv[i] = v;
  dinv32[i] = (float)v;
}

// F_bc: forces with fixed DOFs zeroed (linear_solver.cpp:83).
__global__ void k_bc_rhs(const double* __restrict__ F, const uint8_t* __restrict__ fixcnt,
                         double* __restrict__ Fb, int64_t n_dof) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n_dof) return;
  Fb[i] = fixcnt[i] ? 0.0 : F[i];
}


// ---------------------------------------------------------------------------------------
// PCG (Eigen conjugate_gradient loop with a Jacobi preconditioner)
// ---------------------------------------------------------------------------------------

// z = Dinv r; block partials of |b|^2, |r|^2, r.z (Eigen: rhsNorm2, residualNorm2, absNew).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_init(const double* __restrict__ b,
                                                     const T* __restrict__ r,
                                                     T* __restrict__ z,
                                                     const T* __restrict__ dinv, int64_t n,
                                                     double* partials, int stride) {
  double acc[3] = {0.0, 0.0, 0.0};
  AMT_GRID_STRIDE(i, n) {
    const T rv = r[i];
    const T zv = dinv[i] * rv;
    z[i] = zv;
    if (b) acc[0] = fma(b[i], b[i], acc[0]);
    acc[1] = fma((double)rv, (double)rv, acc[1]);
    acc[2] = fma((double)rv, (double)zv, acc[2]);
  }
  store_partials<3>(acc, partials, stride);
}

// One block: S[0] from the init partials. Without b (the inner solve of a refinement step)
// the tolerance is relative to |r0|.
__global__ __launch_bounds__(kBlock) void k_pcg_start(const double* __restrict__ partials,
                                                      int nb, int stride, int has_b,
                                                      double tol, int maxit, PcgState* S) {
  double t[3];
  reduce_partials<3>(partials, nb, stride, t);
  if (threadIdx.x == 0) {
    PcgState st;
    st.rhs2 = has_b ? t[0] : t[1];
    st.thresh = fmax(tol * tol * st.rhs2, 2.2250738585072014e-308);
    st.rr = t[1];
    st.rz = t[2];
    st.it = 0;
    st.maxit = maxit;
    const int conv = (st.rhs2 == 0.0) || (t[1] < st.thresh);
    st.converged = conv;
    st.done = conv || maxit <= 0;
    S[0] = st;
  }
}

// Eigen loop bookkeeping at the top of iteration t (t > 0): the residual of update(t-1)
// decides convergence (iterations stay t-1... counted as cur.it), else beta and it+1.
// Returns the state iteration t continues with; beta is set when not done.
__device__ __forceinline__ PcgState pcg_step_state(const PcgState& cur, int t,
                                                   const double* upart, int nbu, int stride,
                                                   double& beta) {
  PcgState nx = cur;
  beta = 0.0;
  if (t == 0) return nx;
  double v[2];
  reduce_partials<2>(upart, nbu, stride, v);
  nx.rr = v[0];
  if (v[0] < cur.thresh) {
    nx.done = 1;
    nx.converged = 1;
    return nx;
  }
  beta = v[1] / cur.rz;
  nx.rz = v[1];
  nx.it = cur.it + 1;
  if (nx.it >= cur.maxit) nx.done = 1;
  return nx;
}

// Iteration t, first half: p = z + beta p_old (formed once per node in the LDS tile),
// q = K p, partials p.q (one per block of the tile grid).
template <typename T, typename TK>
__global__ __launch_bounds__(kBlock) void k_pcg_apply(
    int t, const T* __restrict__ z, const T* __restrict__ pold, T* __restrict__ pnew,
    T* __restrict__ q, const T* __restrict__ s, const TK* __restrict__ Ke,
    const uint8_t* __restrict__ fixcnt, const double* __restrict__ P, GridDims g, TileGrid tg,
    PcgState* S, const double* __restrict__ upart, int nbu, double* __restrict__ apart,
    int stride) {
  const PcgState cur = S[t & 1];
  if (cur.done) {
    if (blockIdx.x == 0 && threadIdx.x == 0) S[(t + 1) & 1] = cur;
    return;
  }
  double beta_d;
  const PcgState nx = pcg_step_state(cur, t, upart, nbu, stride, beta_d);
  if (blockIdx.x == 0 && threadIdx.x == 0) S[(t + 1) & 1] = nx;
  if (nx.done) return;
  const T beta = (T)beta_d;
  __shared__ TileSmem<T> sm;
  int i0, j0, k0;
  const bool has = tile_origin(tg, i0, j0, k0);
  if (has)
    tile_load(sm, g, i0, j0, k0, [&](int64_t qq) { return fma(beta, pold[qq], z[qq]); }, s);
  __syncthreads();
  int li, lj, lk;
  tile_thread(li, lj, lk);
  const int i = i0 + li, j = j0 + lj, k = k0 + lk;
  double acc[1] = {0.0};
  if (has && i <= g.nx && j <= g.ny && k <= g.nz) {
    const int64_t n = ((int64_t)i * g.nny + j) * g.nnz + k;
    Nbr<T> nb;
    T s8[8];
    tile_fetch(sm, li, lj, lk, nb, s8);
    T yv[3];
    apply_node(nb, s8, Ke, yv);
    const T pen = (T)(*P);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      yv[c] = fma((T)fixcnt[3 * n + c] * pen, nb.v[13][c], yv[c]);
      pnew[3 * n + c] = nb.v[13][c];
      q[3 * n + c] = yv[c];
      acc[0] = fma((double)nb.v[13][c], (double)yv[c], acc[0]);
    }
  }
  store_partials<1>(acc, apart, stride);
}

// Iteration t, second half: alpha = rz / p.q; x += alpha p; r -= alpha q; z = Dinv r;
// partials |r|^2 and r.z (consumed by iteration t+1 or k_pcg_check).
template <typename T, typename TX>
__global__ __launch_bounds__(kBlock) void k_pcg_update(
    int t, TX* __restrict__ x, T* __restrict__ r, T* __restrict__ z, const T* __restrict__ p,
    const T* __restrict__ q, const T* __restrict__ dinv, int64_t n, const PcgState* S,
    const double* __restrict__ apart, int nba, double* __restrict__ upart, int stride) {
  const PcgState st = S[(t + 1) & 1];
  if (st.done) return;
  double pq[1];
  reduce_partials<1>(apart, nba, stride, pq);
  const T a = (T)(st.rz / pq[0]);
  double acc[2] = {0.0, 0.0};
  AMT_GRID_STRIDE(i, n) {
    const T pv = p[i];
    x[i] = x[i] + (TX)(a * pv);
    const T rv = r[i] - a * q[i];
    r[i] = rv;
    const T zv = dinv[i] * rv;
    z[i] = zv;
    acc[0] = fma((double)rv, (double)rv, acc[0]);
    acc[1] = fma((double)rv, (double)zv, acc[1]);
  }
  store_partials<2>(acc, upart, stride);
}

// One block: the bookkeeping iteration t would do, written to *flag for the host to poll.
__global__ __launch_bounds__(kBlock) void k_pcg_check(int t, const PcgState* S,
                                                      const double* __restrict__ upart,
                                                      int nbu, int stride, PcgState* flag) {
  const PcgState cur = S[t & 1];
  if (cur.done) {
    if (threadIdx.x == 0) *flag = cur;
    return;
  }
  double beta;
  const PcgState nx = pcg_step_state(cur, t, upart, nbu, stride, beta);
  if (threadIdx.x == 0) *flag = nx;
}

// ---------------------------------------------------------------------------------------
// Dot product (compliance u.F, topology_optimizer.hpp:273-278)
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_dot(const double* __restrict__ a,
                                                const double* __restrict__ b, int64_t n,
                                                double* partials, int stride) {
  double acc[1] = {0.0};
  AMT_GRID_STRIDE(i, n) acc[0] = fma(a[i], b[i], acc[0]);
  store_partials<1>(acc, partials, stride);
}

// ---------------------------------------------------------------------------------------
// Warm start in span{u1, u2} (u1 the previous displacement, u2 the one before): the start
// that minimises the energy 1/2 u^T K u - F.u over the span. With r = F - K u1 and w = K u2,
// the 2x2 system needs u1.F, u1.r, u2.F, u2.r and u2.w only, and the new start's residual is
// (1 - c1) F + c1 r - c2 w: one operator application (K u2) in all. It is never worse in
// energy than u1 alone; a degenerate system (u2 = 0 or parallel to u1) falls back to the
// energy-optimal multiple of u1.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_proj_dots(const double* __restrict__ u1,
                                                      const double* __restrict__ u2,
                                                      const double* __restrict__ F,
                                                      const double* __restrict__ r,
                                                      const double* __restrict__ w, int64_t n,
                                                      double* partials, int stride) {
  double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  AMT_GRID_STRIDE(i, n) {
    const double a = u1[i], b = u2[i], f = F[i], rr = r[i];
    v[0] = fma(a, f, v[0]);
    v[1] = fma(a, rr, v[1]);
    v[2] = fma(b, f, v[2]);
    v[3] = fma(b, rr, v[3]);
    v[4] = fma(b, w[i], v[4]);
  }
  store_partials_many<5>(v, partials, stride);
}

// c = argmin of the energy over span{u1, u2} (one block); out[0..1] = c1, c2.
__global__ __launch_bounds__(kBlock) void k_proj_coef(const double* partials, int nb, int stride,
                                                      double* out) {
  double t[5];
  reduce_partials<5>(partials, nb, stride, t);
  if (threadIdx.x) return;
  const double a11 = t[0] - t[1], a12 = t[2] - t[3], a22 = t[4], b1 = t[0], b2 = t[2];
  const double det = a11 * a22 - a12 * a12;
  double c1 = 1.0, c2 = 0.0;
  if (a11 > 0.0 && a22 > 0.0 && det > 1e-10 * a11 * a22) {
    c1 = (b1 * a22 - b2 * a12) / det;
    c2 = (a11 * b2 - a12 * b1) / det;
  } else if (a11 > 0.0) {
    c1 = b1 / a11;
  }
  if (!isfinite(c1) || !isfinite(c2)) {
    c1 = 1.0;
    c2 = 0.0;
  }
  out[0] = c1;
  out[1] = c2;
}

// u <- c1 u + c2 u2, u2 <- u (the old value), r <- (1 - c1) F + c1 r - c2 w; partials of |r|^2.
__global__ __launch_bounds__(kBlock) void k_proj_apply(double* __restrict__ u,
                                                       double* __restrict__ u2,
                                                       const double* __restrict__ F,
                                                       double* __restrict__ r,
                                                       const double* __restrict__ w, int64_t n,
                                                       const double* __restrict__ c,
                                                       double* partials, int stride) {
  const double c1 = c[0], c2 = c[1];
  double acc[1] = {0.0};
  AMT_GRID_STRIDE(i, n) {
    const double a = u[i];
    u[i] = fma(c1, a, c2 * u2[i]);
    u2[i] = a;
    const double rn = fma(1.0 - c1, F[i], fma(c1, r[i], -c2 * w[i]));
    r[i] = rn;
    acc[0] = fma(rn, rn, acc[0]);
  }
  store_partials<1>(acc, partials, stride);
}

// ---------------------------------------------------------------------------------------
// Sensitivities (topology_optimizer.cpp:126-161): per slot e (x-fastest decode),
// ce = u_e^T Ke0 u_e, dc = -p rho^(p-1) ce.
// ---------------------------------------------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_sens(const double* __restrict__ u,
                                                 const double* __restrict__ rho,
                                                 const double* __restrict__ Ke, GridDims g,
                                                 double penal, double* __restrict__ dc) {
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= g.n_elem) return;
  const int64_t nxy = (int64_t)g.nx * g.ny;
  const int iz = (int)(e / nxy);
  const int64_t rem = e % nxy;
  const int iy = (int)(rem / g.nx), ix = (int)(rem % g.nx);
  double ue[24];
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const int64_t node = ((int64_t)(ix + kOX[b]) * g.nny + (iy + kOY[b])) * g.nnz + (iz + kOZ[b]);
    ue[3 * b + 0] = u[3 * node + 0];
    ue[3 * b + 1] = u[3 * node + 1];
    ue[3 * b + 2] = u[3 * node + 2];
  }
  double ce = 0.0;
#pragma unroll
  for (int b = 0; b < 24; ++b) {
    double t = 0.0;
#pragma unroll
    for (int a = 0; a < 24; ++a) t = fma(ue[a], Ke[a * 24 + b], t);
    ce = fma(t, ue[b], ce);
  }
  dc[e] = -penal * pow(rho[e], penal - 1.0) * ce;
}

// ---------------------------------------------------------------------------------------
// Sensitivity filter (topology_optimizer.cpp:60-119 lists, :168-198 apply).
// The neighbour list of slot e is the interior stencil table (built on the host in
// dz, dy, dx order with the reference's sqrt/weight arithmetic) with out-of-grid entries
// skipped, which is exactly the list buildFilter stores.
// ---------------------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_filter(const double* __restrict__ dc,
                                                   const double* __restrict__ rho,
                                                   const StencilEntry* __restrict__ st,
                                                   const double* __restrict__ w, int n_st,
                                                   GridDims g, double* __restrict__ dcf) {
#pragma clang fp contract(off)
  const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (e >= g.n_elem) return;
  const int64_t nxy = (int64_t)g.nx * g.ny;
  const int iz = (int)(e / nxy);
  const int64_t rem = e % nxy;
  const int iy = (int)(rem / g.nx), ix = (int)(rem % g.nx);
  double num = 0.0, den = 0.0;
  for (int t = 0; t < n_st; ++t) {
    const StencilEntry o = st[t];
    const int cx = ix + o.dx, cy = iy + o.dy, cz = iz + o.dz;
    if (cx < 0 || cx >= g.nx || cy < 0 || cy >= g.ny || cz < 0 || cz >= g.nz) continue;
    const int64_t m = (int64_t)cz * nxy + (int64_t)cy * g.nx + cx;
    const double wr = w[t] * rho[m];
    num = num + wr * dc[m];
    den = den + wr;
  }
  dcf[e] = den > 1e-15 ? num / den : dc[e];
}

// LDS-tiled sensitivity filter: k_filter's operation, bit for bit. A block owns a 32 x ty x tz
// element tile (x, the slot-fastest axis, along the threads) and stages rho and dc of the tile and
// a halo of R elements (R = the stencil's reach, <= 3) into LDS with coalesced loads. Each thread
// then walks the stencil for its elements in the same order, with the same bounds test and the
// same arithmetic (no contraction), reading the neighbours from LDS; the stencil table is read
// with block-uniform loads. Tile by reach: R <= 1: 32 x 8 x 8; R = 2: 32 x 8 x 4; R = 3:
// 32 x 4 x 4 (halo at most 38 x 10 x 10 = 3,800 elements, 61 KB).
constexpr int kFilterTX = 32;
constexpr int kFilterLds = 3840;
inline void filter_tile_dims(int R, int& ty, int& tz) {
  ty = R <= 2 ? 8 : 4;
  tz = R <= 1 ? 8 : 4;
}
__global__ __launch_bounds__(kBlock) void k_filter_tile(const double* __restrict__ dc,
                                                        const double* __restrict__ rho,
                                                        const StencilEntry* __restrict__ st,
                                                        const double* __restrict__ w, int n_st,
                                                        GridDims g, int R, int ty, int tz,
                                                        double* __restrict__ dcf) {
#pragma clang fp contract(off)
  __shared__ double srho[kFilterLds], sdc[kFilterLds];
  const int hx = kFilterTX + 2 * R, hy = ty + 2 * R, hz = tz + 2 * R;
  const int ntx = (g.nx + kFilterTX - 1) / kFilterTX, nty = (g.ny + ty - 1) / ty;
  const int b = blockIdx.x;
  const int x0 = (b % ntx) * kFilterTX, y0 = ((b / ntx) % nty) * ty, z0 = (b / (ntx * nty)) * tz;
  const int64_t nxy = (int64_t)g.nx * g.ny;
  for (int q = threadIdx.x; q < hx * hy * hz; q += kBlock) {
    const int gx = x0 - R + q % hx, gy = y0 - R + (q / hx) % hy, gz = z0 - R + q / (hx * hy);
    const bool in = gx >= 0 && gx < g.nx && gy >= 0 && gy < g.ny && gz >= 0 && gz < g.nz;
    const int64_t m = in ? (int64_t)gz * nxy + (int64_t)gy * g.nx + gx : 0;
    srho[q] = in ? rho[m] : 0.0;  // halo cells outside the grid are never read
    sdc[q] = in ? dc[m] : 0.0;
  }
  __syncthreads();
  for (int le = threadIdx.x; le < kFilterTX * ty * tz; le += kBlock) {
    const int lx = le % kFilterTX, ly = (le / kFilterTX) % ty, lz = le / (kFilterTX * ty);
    const int ix = x0 + lx, iy = y0 + ly, iz = z0 + lz;
    if (ix >= g.nx || iy >= g.ny || iz >= g.nz) continue;
    const int c0 = ((lz + R) * hy + (ly + R)) * hx + (lx + R);
    double num = 0.0, den = 0.0;
    for (int t = 0; t < n_st; ++t) {
      const StencilEntry o = st[t];
      const int cx = ix + o.dx, cy = iy + o.dy, cz = iz + o.dz;
      if (cx < 0 || cx >= g.nx || cy < 0 || cy >= g.ny || cz < 0 || cz >= g.nz) continue;
      const int h = c0 + (o.dz * hy + o.dy) * hx + o.dx;
      const double wr = w[t] * srho[h];
      num = num + wr * sdc[h];
      den = den + wr;
    }
    dcf[(int64_t)iz * nxy + (int64_t)iy * g.nx + ix] = den > 1e-15 ? num / den : sdc[c0];
  }
}

// Writes the neighbour list of slot e exactly as k_filter walks it.
__global__ void k_filter_list(int64_t e, const StencilEntry* __restrict__ st,
                              const double* __restrict__ w, int n_st, GridDims g,
                              int64_t* __restrict__ idx, double* __restrict__ wout, int cap,
                              int* count) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int64_t nxy = (int64_t)g.nx * g.ny;
  const int iz = (int)(e / nxy);
  const int64_t rem = e % nxy;
  const int iy = (int)(rem / g.nx), ix = (int)(rem % g.nx);
  int c = 0;
  for (int t = 0; t < n_st; ++t) {
    const StencilEntry o = st[t];
    const int cx = ix + o.dx, cy = iy + o.dy, cz = iz + o.dz;
    if (cx < 0 || cx >= g.nx || cy < 0 || cy >= g.ny || cz < 0 || cz >= g.nz) continue;
    if (c < cap) {
      idx[c] = (int64_t)cz * nxy + (int64_t)cy * g.nx + cx;
      wout[c] = w[t];
    }
    ++c;
  }
  *count = c;
}

// ---------------------------------------------------------------------------------------
// Optimality Criteria (topology_optimizer.cpp:205-287)
//
// The reference runs up to 100 dependent bisection passes, each a full reduction. Here one
// launch evaluates the next LEV levels of the bisection tree at once: the 2^LEV-1 midpoints
// every path can visit, computed with the same (l1+l2)/2 arithmetic. The last block then
// walks the tree with the reference's break test, so the visited lambdas and the l1/l2
// updates are the reference's, in ~LEV times fewer grid-wide reductions.
// ---------------------------------------------------------------------------------------

// Candidate volumes for the 2^LEV - 1 midpoints of the next LEV bisection levels. Every
// candidate is the reference's own expression, rho * sqrt(-dc / (lambda + 1e-20)) clamped
// (topology_optimizer.cpp:249-256), with IEEE division and square root (correctly rounded on
// the device as on the host), so each element's candidate is bitwise the reference's; only the
// order of the volume sum differs (a fixed tree here, an OpenMP static reduction there).
template <int LEV>
__global__ __launch_bounds__(kBlock) void k_oc_sweep(const double* __restrict__ rho,
                                                     const double* __restrict__ dcf,
                                                     const uint8_t* __restrict__ design,
                                                     int64_t n, double move, double rho_min,
                                                     const OcState* st, double* partials,
                                                     int stride) {
  if (st->done) return;
  constexpr int NC = (1 << LEV) - 1;
  __shared__ double lmid[NC];
  if (threadIdx.x < NC) lmid[threadIdx.x] = oc_tree_mid(st->l1, st->l2, threadIdx.x + 1);
  __syncthreads();
  double vol[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) vol[c] = 0.0;
  AMT_GRID_STRIDE(e, n) {
    if (!design[e]) continue;
    const double r = rho[e], d = dcf[e];
#pragma unroll
    for (int c = 0; c < NC; ++c) vol[c] += oc_candidate(r, d, lmid[c], move, rho_min);
  }
  store_partials_many<NC>(vol, partials, stride);
}

// Bisection state at the start of an OC update (topology_optimizer.cpp:215-225): [1e-20, 1e20],
// the volume target, and the design-element count (all elements when the mask is empty).
__global__ void k_oc_init(OcState* st, double vf, const double* __restrict__ n_design,
                          int64_t n_e) {
  if (threadIdx.x || blockIdx.x) return;
  OcState s;
  memset(&s, 0, sizeof(s));
  s.l1 = 1e-20;
  s.l2 = 1e20;
  s.vf = vf;
  s.n_design = *n_design > 0 ? *n_design : (double)n_e;
  *st = s;
}

// End-of-iteration scalars gathered for one host read: OC state, compliance, |b - K u|^2.
struct IterScalars {
  OcState oc;
  double compl_, rr;
};
__global__ void k_iter_scalars(const OcState* __restrict__ oc, const double* __restrict__ compl_,
                               const double* __restrict__ rr, IterScalars* __restrict__ out) {
  if (threadIdx.x) return;
  out->oc = *oc;
  out->compl_ = *compl_;
  out->rr = *rr;
}

// Bisection passes with a known outcome. A design element with dc < 0 sits at its floor
// max(rho_min, rho - move) once rho sqrt(-dc / (lambda + 1e-20)) <= that floor, i.e. for
// lambda + 1e-20 >= rho^2 (-dc) / floor^2, and at its ceiling once the candidate reaches
// rho + move, i.e. for lambda + 1e-20 <= rho^2 (-dc) / (rho + move)^2; elements with dc >= 0
// take the same value at every lambda. Beyond the largest of the first bounds (below the
// smallest of the second) the design volume is therefore a constant, computed once, and the
// pass's l1/l2 decision is the one evaluating it would give. The bisection starts at
// [1e-20, 1e20], so most of its ~80 passes only move l2 down to this range; they are applied
// here in a few instructions instead of being evaluated sweep by sweep.
__device__ __forceinline__ void oc_fast_forward(OcState& s) {
  while (!s.done) {
    if (s.passes >= 100 || (s.l2 - s.l1) / (s.l1 + s.l2) < 1e-4) {
      s.done = 1;
      break;
    }
    const double lm = 0.5 * (s.l1 + s.l2);
    double vol;
    if (lm + 1e-20 >= s.ff_hi)
      vol = s.vol_hi;
    else if (lm + 1e-20 <= s.ff_lo)
      vol = s.vol_lo;
    else
      break;
    s.lmid = lm;
    ++s.passes;
    if (vol / s.n_design > s.vf)
      s.l1 = lm;
    else
      s.l2 = lm;
  }
}

// Partials of the fast-forward bounds and volumes (grid and element order of k_oc_sweep):
// [0] sum of candidates at lambda -> infinity, [1] at lambda = 0, [2] max floor bound,
// [3] max of minus the ceiling bound.
__global__ __launch_bounds__(kBlock) void k_oc_bounds(const double* __restrict__ rho,
                                                      const double* __restrict__ dcf,
                                                      const uint8_t* __restrict__ design,
                                                      int64_t n, double move, double rho_min,
                                                      double* partials, int stride) {
  double v[2] = {0.0, 0.0}, b[2] = {0.0, -INFINITY};
  AMT_GRID_STRIDE(e, n) {
    if (!design[e]) continue;
    const double r = rho[e], d = dcf[e];
    v[0] += oc_candidate(r, d, 1e300, move, rho_min);
    v[1] += oc_candidate(r, d, 0.0, move, rho_min);
    if (d < 0.0) {
      const double f = std_max(rho_min, r - move), c = r + move;
      b[0] = fmax(b[0], r * r * (-d) / (f * f));
      b[1] = fmax(b[1], -(r * r * (-d) / (c * c)));
    }
  }
  store_partials<2>(v, partials, stride);
  store_partials<2, true>(b, partials + 2 * (size_t)stride, stride);
}

// One block: the bounds into the state (with a 1e-9 relative margin over rounding), then the
// first run of known passes.
__global__ __launch_bounds__(kBlock) void k_oc_bounds_final(OcState* st, const double* partials,
                                                            int nb, int stride) {
  double v[2], b[2];
  reduce_partials<2>(partials, nb, stride, v);
  reduce_partials<2, true>(partials + 2 * (size_t)stride, nb, stride, b);
  if (threadIdx.x) return;
  OcState s = *st;
  s.vol_hi = v[0];
  s.vol_lo = v[1];
  s.ff_hi = b[0] * (1.0 + 1e-9);
  s.ff_lo = -b[1] * (1.0 - 1e-9);
  oc_fast_forward(s);
  *st = s;
}

// One block: reduce the candidate volumes and walk the bisection tree with the reference's
// break test; updates l1, l2, the last evaluated lambda and the pass count, then applies the
// passes whose outcome is known (oc_fast_forward).
template <int LEV>
__global__ __launch_bounds__(kBlock) void k_oc_walk(OcState* st,
                                                    const double* __restrict__ partials,
                                                    int nb, int stride) {
  if (st->done) return;
  constexpr int NC = (1 << LEV) - 1;
  __shared__ double tots[NC];
  double tot[NC];
  reduce_partials<NC>(partials, nb, stride, tot);
  if (threadIdx.x != 0) return;
#pragma unroll
  for (int c = 0; c < NC; ++c) tots[c] = tot[c];
  double l1 = st->l1, l2 = st->l2;
  int nd = 1, passes = st->passes, done = 0;
  for (int lev = 0; lev < LEV; ++lev) {
    if (passes >= 100 || (l2 - l1) / (l1 + l2) < 1e-4) {
      done = 1;
      break;
    }
    const double lm = 0.5 * (l1 + l2);  // == oc_tree_mid(st->l1, st->l2, nd)
    st->lmid = lm;
    ++passes;
    if (tots[nd - 1] / st->n_design > st->vf) {
      l1 = lm;
      nd = 2 * nd + 1;
    } else {
      l2 = lm;
      nd = 2 * nd;
    }
  }
  if (!done && (passes >= 100 || (l2 - l1) / (l1 + l2) < 1e-4)) done = 1;
  OcState s = *st;
  s.l1 = l1;
  s.l2 = l2;
  s.passes = passes;
  s.done = done;
  oc_fast_forward(s);
  *st = s;
}

// rho <- candidate at the last evaluated lambda (design) / unchanged (non-design);
// partials of max|drho| and the design-space volume (topology_optimizer.cpp:274-286, :364-381).
__global__ __launch_bounds__(kBlock) void k_oc_final(double* __restrict__ rho,
                                                     const double* __restrict__ dcf,
                                                     const uint8_t* __restrict__ design,
                                                     int64_t n, double move, double rho_min,
                                                     const OcState* st, double* partials,
                                                     int stride) {
  const double lm = st->lmid;
  const int any = st->passes > 0;
  double acc[1] = {0.0}, mx[1] = {0.0};
  AMT_GRID_STRIDE(e, n) {
    const double r = rho[e];
    double rn = r;
    if (design[e]) {
      rn = any ? oc_candidate(r, dcf[e], lm, move, rho_min) : 0.0;
      acc[0] += rn;
    }
    mx[0] = fmax(mx[0], fabs(rn - r));
    rho[e] = rn;
  }
  store_partials<1, true>(mx, partials, stride);
  store_partials<1>(acc, partials + stride, stride);
}

__global__ __launch_bounds__(kBlock) void k_oc_final_reduce(OcState* st,
                                                            const double* __restrict__ partials,
                                                            int nb, int stride) {
  double m[1], v[1];
  reduce_partials<1, true>(partials, nb, stride, m);
  reduce_partials<1>(partials + stride, nb, stride, v);
  if (threadIdx.x == 0) {
    st->change = m[0];
    st->vol_sum = v[0];
  }
}

// Partial counts of a caller-supplied design mask (the OC update on caller buffers).
__global__ __launch_bounds__(kBlock) void k_design_count(const uint8_t* __restrict__ design,
                                                         int64_t n, double* partials,
                                                         int stride) {
  double acc[1] = {0.0};
  AMT_GRID_STRIDE(e, n) acc[0] += design[e] ? 1.0 : 0.0;
  store_partials<1>(acc, partials, stride);
}

// One block: the bisection's starting state [1e-20, 1e20] with the design count reduced from
// k_design_count's partials (an empty mask counts every element, topology_optimizer.cpp:222-225).
// Every field is written, so nothing of a previous call survives in the state.
__global__ __launch_bounds__(kBlock) void k_oc_state_init(OcState* st, const double* partials,
                                                          int nb, int stride, int64_t n,
                                                          double vf) {
  double c[1];
  reduce_partials<1>(partials, nb, stride, c);
  if (threadIdx.x) return;
  OcState s;
  s.l1 = 1e-20;
  s.l2 = 1e20;
  s.lmid = 0.0;
  s.vf = vf;
  s.n_design = c[0] > 0.0 ? c[0] : (double)n;
  s.change = 0.0;
  s.vol_sum = 0.0;
  s.pad = 0.0;
  s.passes = 0;
  s.done = 0;
  s.pad1 = s.pad2 = 0;
  s.ff_hi = INFINITY;
  s.ff_lo = -INFINITY;
  s.vol_hi = s.vol_lo = 0.0;
  *st = s;
}

// out[0] = max |drho|, out[1] = design volume sum, out[2] = bisection passes, out[3] = design
// element count.
__global__ void k_oc_state_out(const OcState* __restrict__ st, double* __restrict__ out) {
  if (threadIdx.x || blockIdx.x) return;
  out[0] = st->change;
  out[1] = st->vol_sum;
  out[2] = (double)st->passes;
  out[3] = st->n_design;
}

// Design mask from the initial density (topology_optimizer.cpp:419-428); partial counts.
__global__ __launch_bounds__(kBlock) void k_design_mask(const double* __restrict__ rho,
                                                        uint8_t* __restrict__ design,
                                                        int64_t n, double rho_min,
                                                        double* partials, int stride) {
  double acc[1] = {0.0};
  AMT_GRID_STRIDE(e, n) {
    const double r = rho[e];
    const uint8_t d = (r > rho_min + 1e-6) && (r < 0.99);
    design[e] = d;
    acc[0] += d;
  }
  store_partials<1>(acc, partials, stride);
}

}  // namespace amt
