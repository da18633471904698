// multigrid.hpp — geometric multigrid V-cycle (FP32) used as the preconditioner of a flexible
// CG whose vectors, operator and inner products stay FP64.
//
// Hierarchy: level 0 is the fine voxel grid, applied matrix-free (s_e * Ke0). Level l+1 has
// ceil(e/2) elements per axis of level l; the prolongation P is trilinear (1, 1/2 weights per
// axis; a last odd fine node interpolates from a coarse node one layer past the fine grid).
// Coarse operators are Galerkin, A_{l+1} = P^T A_l P, 27-point block stencils. They are
// symmetric, A(N, N+d) = A(N+d, N)^T, so only the diagonal block and the 13 forward offsets
// d = 13..26 are stored, in SoA layout A[((d-13)*9 + 3i + j) * n_nodes + node] (i/j =
// components); a backward block is read transposed from the neighbour (stencil_coef). This
// halves the coarse levels' traffic and makes the stored operator exactly symmetric:
//  - level 1 straight from the element densities: a coarse element E holds 8 fine children c,
//    and P restricted to E is the fixed interpolation Q_c, so P^T A_0 P over E equals
//    sum_c s_c G_c with G_c = Q_c^T Ke0 Q_c (eight 24x24 tables built on the host); the
//    penalty rows of the fine operator add P^T diag(pen) P;
//  - level l >= 2 by the generic stencil triple product.
// The coarsest level (at most 192 DOFs) is solved with a dense inverse computed on the device.
// Smoothing is one damped-Jacobi sweep before and after the coarse correction (pre-smoother
// started from zero), so the V-cycle is a symmetric linear operator.
#pragma once

#include "kernels.hpp"

namespace amt {

struct MgGrid {
  int ex, ey, ez;  // elements per axis
  int nnx, nny, nnz;
  int64_t n_nodes;
};

__device__ __forceinline__ void mg_decode(int64_t n, const MgGrid& g, int& i, int& j, int& k) {
  k = (int)(n % g.nnz);
  const int64_t t = n / g.nnz;
  j = (int)(t % g.nny);
  i = (int)(t / g.nny);
}
__device__ __forceinline__ int64_t mg_node(const MgGrid& g, int i, int j, int k) {
  return ((int64_t)i * g.nny + j) * g.nnz + k;
}
// 1D trilinear prolongation weight of fine node f from coarse node c
__device__ __forceinline__ float pw1(int f, int c) {
  const int d = f - 2 * c;
  return d == 0 ? 1.0f : ((d == 1 || d == -1) ? 0.5f : 0.0f);
}

constexpr int kStencilBlocks = 14;  // stored 3x3 blocks per node: d = 13..26

// Penalty rows on the coarse levels. A clamped fine node on an odd-sized upper boundary is
// interpolated from two coarse nodes, one of them a layer past the grid, so the Galerkin
// penalty term couples them with P/4 entries and pins only their average; the difference
// keeps its O(K_ii) stiffness. In FP32 that stiffness is below the rounding of the P-sized
// entries (P = 1e8 max K_ii), the coarse operator turns indefinite and the flexible CG
// diverges (clamps on lower faces sit on coarse nodes and are unaffected). The coarse levels
// therefore carry the penalty at 1e-5 P (1e3 max K_ii for the reference's factor), which
// still pins the constrained combinations and is resolved in FP32; level 0 keeps the true
// penalty, and the transfers between levels 0 and 1 skip the clamped fine DOFs (their
// residual is not restricted, their value gets no coarse correction), so the smaller coarse
// penalty never acts on them.
constexpr double kMgCoarsePenalty = 1e-5;

// Block (r, c) of A(n, n + d); the neighbour n + d must exist. sj, si: node strides along j, i.
__device__ __forceinline__ float stencil_coef(const float* __restrict__ A, int64_t nn, int64_t n,
                                              int d, int r, int c, int64_t sj, int64_t si) {
  if (d >= 13) return A[(int64_t)((d - 13) * 9 + 3 * r + c) * nn + n];
  const int64_t m = n + (d / 9 - 1) * si + ((d / 3) % 3 - 1) * sj + (d % 3 - 1);
  return A[(int64_t)((13 - d) * 9 + 3 * c + r) * nn + m];
}

// ---------------------------------------------------------------------------------------
// Coarse operator construction (once per OC iteration)
// ---------------------------------------------------------------------------------------

// Level 1, thread (coarse node N, row component i = blockIdx.z, offset D = 13 + blockIdx.y).
// G: 8 x 576 doubles (child c, row 3a+i, col 3b+j). s: fine element scales, C order.
__global__ __launch_bounds__(kBlock) void k_mg_galerkin1(const double* __restrict__ s,
                                                         GridDims fg,
                                                         const double* __restrict__ G,
                                                         const uint8_t* __restrict__ fixcnt,
                                                         const double* __restrict__ P,
                                                         MgGrid cg, float* __restrict__ A) {
  const int64_t N = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (N >= cg.n_nodes) return;
  const int D = 13 + (int)blockIdx.y;
  const int dI = D / 9 - 1, dJ = (D / 3) % 3 - 1, dK = D % 3 - 1;
  const int i = blockIdx.z;
  int I, J, K;
  mg_decode(N, cg, I, J, K);
  const int MI = I + dI, MJ = J + dJ, MK = K + dK;
  double acc[3] = {0.0, 0.0, 0.0};
  if (MI >= 0 && MI < cg.nnx && MJ >= 0 && MJ < cg.nny && MK >= 0 && MK < cg.nnz) {
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int bx = dI + kOX[a], by = dJ + kOY[a], bz = dK + kOZ[a];
      if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;  // uniform
      int bl = 0;  // local index of M in E (element_kernel.cpp:196-199 node order)
#pragma unroll
      for (int t = 0; t < 8; ++t)
        if (kOX[t] == bx && kOY[t] == by && kOZ[t] == bz) bl = t;
      const int EI = I - kOX[a], EJ = J - kOY[a], EK = K - kOZ[a];
      if (EI < 0 || EI >= cg.ex || EJ < 0 || EJ >= cg.ey || EK < 0 || EK >= cg.ez) continue;
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int fi = 2 * EI + kOX[c], fj = 2 * EJ + kOY[c], fk = 2 * EK + kOZ[c];
        if (fi >= fg.nx || fj >= fg.ny || fk >= fg.nz) continue;
        const double se = s[((int64_t)fi * fg.ny + fj) * fg.nz + fk];
        const double* Gr = G + c * 576 + (3 * a + i) * 24 + 3 * bl;
        acc[0] = fma(se, Gr[0], acc[0]);
        acc[1] = fma(se, Gr[1], acc[1]);
        acc[2] = fma(se, Gr[2], acc[2]);
      }
    }
    // penalty rows: sum_m w_N(m) w_M(m) pen(m,i) on the diagonal component, with the
    // penalty scaled down to kMgCoarsePenalty * P (see the note at kMgCoarsePenalty)
    const double pen = kMgCoarsePenalty * (*P);
    for (int mx = 2 * I - 1; mx <= 2 * I + 1; ++mx) {
      if (mx < 0 || mx > fg.nx || pw1(mx, MI) == 0.0f) continue;
      for (int my = 2 * J - 1; my <= 2 * J + 1; ++my) {
        if (my < 0 || my > fg.ny || pw1(my, MJ) == 0.0f) continue;
        for (int mz = 2 * K - 1; mz <= 2 * K + 1; ++mz) {
          if (mz < 0 || mz > fg.nz || pw1(mz, MK) == 0.0f) continue;
          const int64_t m = ((int64_t)mx * fg.nny + my) * fg.nnz + mz;
          const int cnt = fixcnt[3 * m + i];
          if (!cnt) continue;
          const double w = (double)pw1(mx, I) * pw1(my, J) * pw1(mz, K) * pw1(mx, MI) *
                           pw1(my, MJ) * pw1(mz, MK);
          acc[i] = fma(w, cnt * pen, acc[i]);
        }
      }
    }
  }
  const int d = blockIdx.y;  // stored block index D - 13
#pragma unroll
  for (int j = 0; j < 3; ++j) A[(int64_t)(d * 9 + 3 * i + j) * cg.n_nodes + N] = (float)acc[j];
}

// Level l+1 from level l, one stored block: A_c(N, M)[i][:] with M = N + (DI, DJ, DK) is
// sum over fine m = 2N + a and m' = 2M + b, |m' - m| <= 1 per axis, of
// w(a) w(b) A_f(m, m')[i][:], with w = 1 on the coarse node's own fine node and 1/2 beside it.
// The (a, b) pairs with |2 D + b - a| <= 1 are fixed by the offset, so every term is unrolled at
// compile time; fine nodes outside the grid (and a neighbour M outside the coarse grid) are read
// at a clamped position with weight 0, so all loads issue without a branch. The terms are
// summed in the order m (x, y, z), then m' (x, y, z).
template <int DI, int DJ, int DK, int AX, int AY>
__device__ __forceinline__ void coarsen_block(const float* __restrict__ Af, const MgGrid& fg,
                                              const MgGrid& cg, int I, int J, int K, int i,
                                              float (&acc)[3]) {
  const int64_t nn = fg.n_nodes;
  const int MI = I + DI, MJ = J + DJ, MK = K + DK;
  const bool okM = MI >= 0 && MI < cg.nnx && MJ >= 0 && MJ < cg.nny && MK >= 0 && MK < cg.nnz;
  {
    constexpr int ax = AX, ay = AY;
#pragma unroll
      for (int az = -1; az <= 1; ++az) {
        const int mx = 2 * I + ax, my = 2 * J + ay, mz = 2 * K + az;
        const bool okm = mx >= 0 && mx < fg.nnx && my >= 0 && my < fg.nny && mz >= 0 &&
                         mz < fg.nnz;
        const int64_t m = mg_node(fg, min(max(mx, 0), fg.nnx - 1), min(max(my, 0), fg.nny - 1),
                                  min(max(mz, 0), fg.nnz - 1));
        const float wN = (ax ? 0.5f : 1.f) * (ay ? 0.5f : 1.f) * (az ? 0.5f : 1.f);
#pragma unroll
        for (int bx = -1; bx <= 1; ++bx) {
          const int ddx = 2 * DI + bx - ax;
          if (ddx < -1 || ddx > 1) continue;  // resolved at compile time
#pragma unroll
          for (int by = -1; by <= 1; ++by) {
            const int ddy = 2 * DJ + by - ay;
            if (ddy < -1 || ddy > 1) continue;
#pragma unroll
            for (int bz = -1; bz <= 1; ++bz) {
              const int ddz = 2 * DK + bz - az;
              if (ddz < -1 || ddz > 1) continue;
              const int px = mx + ddx, py = my + ddy, pz = mz + ddz;
              const bool ok = okM && okm && px >= 0 && px < fg.nnx && py >= 0 && py < fg.nny &&
                              pz >= 0 && pz < fg.nnz;
              const float w =
                  ok ? wN * (bx ? 0.5f : 1.f) * (by ? 0.5f : 1.f) * (bz ? 0.5f : 1.f) : 0.f;
              const int d = (ddx + 1) * 9 + (ddy + 1) * 3 + (ddz + 1);
              if (d >= 13) {
#pragma unroll
                for (int jj = 0; jj < 3; ++jj)
                  acc[jj] = fmaf(w, Af[(int64_t)((d - 13) * 9 + 3 * i + jj) * nn + m], acc[jj]);
              } else {
                const int64_t pm = mg_node(fg, min(max(px, 0), fg.nnx - 1),
                                           min(max(py, 0), fg.nny - 1),
                                           min(max(pz, 0), fg.nnz - 1));
#pragma unroll
                for (int jj = 0; jj < 3; ++jj)
                  acc[jj] = fmaf(w, Af[(int64_t)((13 - d) * 9 + 3 * jj + i) * nn + pm], acc[jj]);
              }
            }
          }
        }
      }
  }
}

// Level l+1 from level l: block (64 coarse nodes, offset D = 13 + blockIdx.y, row i =
// blockIdx.z), wave w taking the fine nodes 2N + a with (a_x, a_y) = (w / 3 - 1, w % 3 - 1)
// (the offset and the wave's share are uniform in the wave and select the unrolled term list).
// Nine waves per output cut the serial load chain of each thread to a ninth; their partial
// sums are added in LDS in fixed order.
constexpr int kCoarsenNodes = 64;
template <int DI, int DJ, int DK>
__device__ __forceinline__ void coarsen_part(int w, const float* __restrict__ Af,
                                             const MgGrid& fg, const MgGrid& cg, int I, int J,
                                             int K, int i, float (&acc)[3]) {
  switch (w) {
    case 0: coarsen_block<DI, DJ, DK, -1, -1>(Af, fg, cg, I, J, K, i, acc); break;
    case 1: coarsen_block<DI, DJ, DK, -1, 0>(Af, fg, cg, I, J, K, i, acc); break;
    case 2: coarsen_block<DI, DJ, DK, -1, 1>(Af, fg, cg, I, J, K, i, acc); break;
    case 3: coarsen_block<DI, DJ, DK, 0, -1>(Af, fg, cg, I, J, K, i, acc); break;
    case 4: coarsen_block<DI, DJ, DK, 0, 0>(Af, fg, cg, I, J, K, i, acc); break;
    case 5: coarsen_block<DI, DJ, DK, 0, 1>(Af, fg, cg, I, J, K, i, acc); break;
    case 6: coarsen_block<DI, DJ, DK, 1, -1>(Af, fg, cg, I, J, K, i, acc); break;
    case 7: coarsen_block<DI, DJ, DK, 1, 0>(Af, fg, cg, I, J, K, i, acc); break;
    default: coarsen_block<DI, DJ, DK, 1, 1>(Af, fg, cg, I, J, K, i, acc); break;
  }
}
__device__ __forceinline__ void coarsen_wave(int D, int w, const float* __restrict__ Af,
                                             const MgGrid& fg, const MgGrid& cg, int I, int J,
                                             int K, int i, float (&acc)[3]) {
  switch (D) {
    case 0: coarsen_part<0, 0, 0>(w, Af, fg, cg, I, J, K, i, acc); break;
    case 1: coarsen_part<0, 0, 1>(w, Af, fg, cg, I, J, K, i, acc); break;
    case 2: coarsen_part<0, 1, -1>(w, Af, fg, cg, I, J, K, i, acc); break;
    case 3: coarsen_part<0, 1, 0>(w, Af, fg, cg, I, J, K, i, acc); break;
    case 4: coarsen_part<0, 1, 1>(w, Af, fg, cg, I, J, K, i, acc); break;
    case 5: coarsen_part<1, -1, -1>(w, Af, fg, cg, I, J, K, i, acc); break;
    case 6: coarsen_part<1, -1, 0>(w, Af, fg, cg, I, J, K, i, acc); break;
    case 7: coarsen_part<1, -1, 1>(w, Af, fg, cg, I, J, K, i, acc); break;
    case 8: coarsen_part<1, 0, -1>(w, Af, fg, cg, I, J, K, i, acc); break;
    case 9: coarsen_part<1, 0, 0>(w, Af, fg, cg, I, J, K, i, acc); break;
    case 10: coarsen_part<1, 0, 1>(w, Af, fg, cg, I, J, K, i, acc); break;
    case 11: coarsen_part<1, 1, -1>(w, Af, fg, cg, I, J, K, i, acc); break;
    case 12: coarsen_part<1, 1, 0>(w, Af, fg, cg, I, J, K, i, acc); break;
    default: coarsen_part<1, 1, 1>(w, Af, fg, cg, I, J, K, i, acc); break;
  }
}

__global__ __launch_bounds__(9 * kCoarsenNodes) void k_mg_coarsen(const float* __restrict__ Af,
                                                                   MgGrid fg, MgGrid cg,
                                                                   float* __restrict__ Ac) {
  __shared__ float part[9][kCoarsenNodes][3];
  const int w = threadIdx.x / kCoarsenNodes, l = threadIdx.x % kCoarsenNodes;
  const int64_t N = (int64_t)blockIdx.x * kCoarsenNodes + l;
  const bool ok = N < cg.n_nodes;
  const int i = blockIdx.z;
  float acc[3] = {0.f, 0.f, 0.f};
  if (ok) {
    int I, J, K;
    mg_decode(N, cg, I, J, K);
    coarsen_wave((int)blockIdx.y, w, Af, fg, cg, I, J, K, i, acc);
  }
#pragma unroll
  for (int j = 0; j < 3; ++j) part[w][l][j] = acc[j];
  __syncthreads();
  if (w || !ok) return;
  const int d = blockIdx.y;  // stored block index D - 13
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    float v = part[0][l][j];
#pragma unroll
    for (int u = 1; u < 9; ++u) v += part[u][l][j];
    Ac[(int64_t)(d * 9 + 3 * i + j) * cg.n_nodes + N] = v;
  }
}

// Smoother weights. Damped Jacobi w_i = omega / A_ii diverges once omega * lambda_max(D^-1 A)
// exceeds 2: for hex8 elasticity lambda_max(D^-1 K) is ~3.1 on a uniform grid but ~4.8 on a
// high-contrast solid/void design (nu = 0.3), so omega = 0.6 can amplify some modes and turn the
// V-cycle into a poor preconditioner. Each row therefore gets
//   w_i = min(omega / A_ii, theta / sum_j |A_ij|),
// which by Gershgorin bounds the spectrum of W A by theta < 2 (a convergent, symmetric
// smoother) while rows with a tame neighbourhood keep close to omega. 0 on empty rows.
__device__ __forceinline__ float smoother_weight(double diag, double rowabs, float omega,
                                                 float theta) {
  if (!(diag > 0.0)) return 0.f;
  const double w = fmin((double)omega / diag, (double)theta / rowabs);
  return (float)w;
}

// Level 0, thread (node, row r = blockIdx.y): the assembled row of K (the 27 neighbour blocks
// summed from the node's 8 elements, plus the penalty on the diagonal) gives A_rr and
// sum_j |A_rj|.
__global__ __launch_bounds__(kBlock) void k_mg_wd0(const double* __restrict__ s,
                                                   const double* __restrict__ Ke, GridDims g,
                                                   const uint8_t* __restrict__ fixcnt,
                                                   const double* __restrict__ P, float omega,
                                                   float theta, float* __restrict__ wd) {
  const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.n_nodes) return;
  const int r = blockIdx.y;
  int i, j, k;
  node_decode(n, g, i, j, k);
  double s8[8];
  load_element_scales(s, g, i, j, k, s8);
  // neighbour block d = (di, dj, dk) gathers the element pairs (a, b) with b - a = d; unrolled,
  // so the pair tests and the Ke0 offsets are compile-time (64 pairs of 3 FMAs)
  double rowabs = 0.0, diag = 0.0;
#pragma unroll
  for (int d = 0; d < 27; ++d) {
    const int di = d / 9 - 1, dj = (d / 3) % 3 - 1, dk = d % 3 - 1;
    double blk[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int a = 0; a < 8; ++a) {
      const int bx = kOX[a] + di, by = kOY[a] + dj, bz = kOZ[a] + dk;
      if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
      const int b = 4 * bz + 2 * by + (bx ^ by);  // corner at offsets (bx, by, bz)
#pragma unroll
      for (int c = 0; c < 3; ++c) blk[c] = fma(s8[a], Ke[(3 * a + r) * 24 + 3 * b + c], blk[c]);
    }
    if (d == 13) {
      blk[r] += (double)fixcnt[3 * n + r] * (*P);
      diag = blk[r];
    }
    rowabs += fabs(blk[0]) + fabs(blk[1]) + fabs(blk[2]);
  }
  wd[3 * n + r] = smoother_weight(diag, rowabs, omega, theta);
}

// Stencil level, thread per (node, row): the same weights from the stored stencil. With mi
// (a Chebyshev level), mi = 1 / sum_j |A_ij| (the preconditioner of the Chebyshev smoother,
// under which the spectrum of M^-1 A lies in (0, 1] by Gershgorin) and wd = mi * w1, the
// smoother's first step from zero; both 0 on rows without a positive diagonal.
__global__ void k_mg_stencil_wdinv(const float* __restrict__ A, MgGrid g, float omega,
                                   float theta, float* __restrict__ wd,
                                   float* __restrict__ mi = nullptr, float w1 = 0.f) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= 3 * g.n_nodes) return;
  const int64_t n = t / 3;
  const int r = (int)(t % 3);
  int i, j, k;
  mg_decode(n, g, i, j, k);
  const int64_t sj = g.nnz, si = (int64_t)g.nny * g.nnz;
  double rowabs = 0.0;
#pragma unroll
  for (int d = 0; d < 27; ++d) {
    const int di = d / 9 - 1, dj = (d / 3) % 3 - 1, dk = d % 3 - 1;
    if ((i + di) < 0 || (i + di) >= g.nnx || (j + dj) < 0 || (j + dj) >= g.nny || (k + dk) < 0 ||
        (k + dk) >= g.nnz)
      continue;
#pragma unroll
    for (int c = 0; c < 3; ++c) rowabs += fabs((double)stencil_coef(A, g.n_nodes, n, d, r, c, sj, si));
  }
  const double diag = (double)A[(int64_t)(4 * r) * g.n_nodes + n];
  if (mi) {
    const float m = (diag > 0.0 && rowabs > 0.0) ? (float)(1.0 / rowabs) : 0.f;
    mi[t] = m;
    wd[t] = m * w1;
  } else {
    wd[t] = smoother_weight(diag, rowabs, omega, theta);
  }
}

// ---------------------------------------------------------------------------------------
// V-cycle kernels. The level-0 sweeps return at once when the outer solve is done; the
// coarse-level kernels (k_mg_pre/post/restrict/prolong/dense) touch only the hierarchy's own
// vectors, so they run unconditionally: a chunk enqueued past convergence recomputes scratch
// values nobody reads, and each launch saves the flag load (~1% of an FCG iteration).
// ---------------------------------------------------------------------------------------
// Loads f(q) for the 27 neighbours of node n (0 outside the grid); f maps a DOF to a value.
template <typename F>
__device__ __forceinline__ void mg_gather(const MgGrid& g, int64_t n, int i, int j, int k, F f,
                                          float (&nb)[27][3]) {
  const int64_t sj = g.nnz, si = (int64_t)g.nny * g.nnz;
#pragma unroll
  for (int di = -1; di <= 1; ++di)
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
      for (int dk = -1; dk <= 1; ++dk) {
        const int m = (di + 1) * 9 + (dj + 1) * 3 + (dk + 1);
        const bool ok = (i + di) >= 0 && (i + di) < g.nnx && (j + dj) >= 0 &&
                        (j + dj) < g.nny && (k + dk) >= 0 && (k + dk) < g.nnz;
        // out-of-grid neighbours load their own node (always valid) and are zeroed after the
        // load, so no load sits behind a branch and all 81 can be in flight together
        const int64_t q = 3 * (ok ? n + di * si + dj * sj + dk : n);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const float v = f(q + c);
          nb[m][c] = ok ? v : 0.f;
        }
      }
}

// Row r of the stencil product at node n = (i, j, k): sum_d sum_c A(n, n+d)[r][c] nb(d, c),
// nb zero outside the grid. A backward block whose neighbour is outside the grid is read at n
// instead (finite, times nb = 0): every load is unconditional.
template <typename NB>
__device__ __forceinline__ float stencil_row_f(const float* __restrict__ A, const MgGrid& g,
                                               int64_t n, int i, int j, int k, int r, NB nb) {
  const int64_t sj = g.nnz, si = (int64_t)g.nny * g.nnz, nn = g.n_nodes;
  float y = 0.f;
#pragma unroll
  for (int d = 0; d < 27; ++d) {
    const int di = d / 9 - 1, dj = (d / 3) % 3 - 1, dk = d % 3 - 1;
    if (d >= 13) {
#pragma unroll
      for (int c = 0; c < 3; ++c) y = fmaf(A[(int64_t)((d - 13) * 9 + 3 * r + c) * nn + n], nb(d, c), y);
    } else {
      const bool ok = (i + di) >= 0 && (i + di) < g.nnx && (j + dj) >= 0 &&
                      (j + dj) < g.nny && (k + dk) >= 0 && (k + dk) < g.nnz;
      const int64_t m = ok ? n + di * si + dj * sj + dk : n;
#pragma unroll
      for (int c = 0; c < 3; ++c) y = fmaf(A[(int64_t)((13 - d) * 9 + 3 * c + r) * nn + m], nb(d, c), y);
    }
  }
  return y;
}
// The same with the neighbour values gathered in registers (mg_gather).
__device__ __forceinline__ float stencil_row(const float* __restrict__ A, const MgGrid& g,
                                             int64_t n, int i, int j, int k, int r,
                                             const float (&nb)[27][3]) {
  return stencil_row_f(A, g, n, i, j, k, r, [&](int d, int c) { return nb[d][c]; });
}

// Stencil level, row (node n, component r): t = rhs - A x0 with the pre-smoothing step from
// zero, x0 = wd*rhs, already written by the restriction that produced rhs.
__device__ __forceinline__ void mg_pre_row(const float* __restrict__ A, const MgGrid& g,
                                           const float* __restrict__ rhs,
                                           const float* __restrict__ x0, float* __restrict__ t,
                                           int64_t n, int r) {
  int i, j, k;
  mg_decode(n, g, i, j, k);
  float nb[27][3];
  mg_gather(g, n, i, j, k, [&](int64_t q) { return x0[q]; }, nb);
  const float y = stencil_row(A, g, n, i, j, k, r, nb);
  const int64_t q = 3 * n + r;
  t[q] = rhs[q] - y;
}

// Stencil level, row (node n, component r): out = x + wd*(rhs - A x) (post-smoothing).
__device__ __forceinline__ void mg_post_row(const float* __restrict__ A, const MgGrid& g,
                                            const float* __restrict__ wd,
                                            const float* __restrict__ rhs,
                                            const float* __restrict__ x, float* __restrict__ out,
                                            int64_t n, int r) {
  int i, j, k;
  mg_decode(n, g, i, j, k);
  float nb[27][3];
  mg_gather(g, n, i, j, k, [&](int64_t q) { return x[q]; }, nb);
  const float y = stencil_row(A, g, n, i, j, k, r, nb);
  const int64_t q = 3 * n + r;
  out[q] = x[q] + wd[q] * (rhs[q] - y);
}

// Half-precision copy of a stencil level's blocks for its smoothing sweeps: every stored entry
// divided by the largest |entry| of its node (S[node]), so each node's 126 values lie in
// [-1, 1] whatever the density around it (void regions sit ~1e-9 below solid ones), and the
// sweep multiplies each block's sum by its node's scale. It halves the bytes of the level's
// sweeps, which stream the stored stencil once per sweep; the coarsening reads the FP32 blocks.
__global__ __launch_bounds__(kBlock) void k_mg_stencil_f16(const float* __restrict__ A, MgGrid g,
                                                           _Float16* __restrict__ Ah,
                                                           float* __restrict__ S) {
  const int64_t N = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (N >= g.n_nodes) return;
  const int64_t nn = g.n_nodes;
  float mx = 0.f;
  for (int e = 0; e < 126; ++e) mx = fmaxf(mx, fabsf(A[e * nn + N]));
  const float sc = mx > 0.f ? mx : 1.f, inv = 1.f / sc;
  for (int e = 0; e < 126; ++e) Ah[e * nn + N] = (_Float16)(A[e * nn + N] * inv);
  S[N] = sc;
}

// stencil_row with the half-precision blocks: each block's three products summed, then scaled
// by the scale of the node that stores the block.
__device__ __forceinline__ float stencil_row_h(const _Float16* __restrict__ Ah,
                                               const float* __restrict__ S, const MgGrid& g,
                                               int64_t n, int i, int j, int k, int r,
                                               const float (&nb)[27][3]) {
  const int64_t sj = g.nnz, si = (int64_t)g.nny * g.nnz, nn = g.n_nodes;
  float y = 0.f, yo = 0.f;  // yo: the blocks stored at n
#pragma unroll
  for (int d = 0; d < 27; ++d) {
    const int di = d / 9 - 1, dj = (d / 3) % 3 - 1, dk = d % 3 - 1;
    float b = 0.f;
    if (d >= 13) {
#pragma unroll
      for (int c = 0; c < 3; ++c)
        b = fmaf((float)Ah[(int64_t)((d - 13) * 9 + 3 * r + c) * nn + n], nb[d][c], b);
      yo += b;
    } else {
      const bool ok = (i + di) >= 0 && (i + di) < g.nnx && (j + dj) >= 0 &&
                      (j + dj) < g.nny && (k + dk) >= 0 && (k + dk) < g.nnz;
      const int64_t m = ok ? n + di * si + dj * sj + dk : n;
#pragma unroll
      for (int c = 0; c < 3; ++c)
        b = fmaf((float)Ah[(int64_t)((13 - d) * 9 + 3 * c + r) * nn + m], nb[d][c], b);
      y = fmaf(S[m], b, y);
    }
  }
  return fmaf(S[n], yo, y);
}

// Restriction sum (P^T r_f)(N) at coarse node N = (I, J, K), r_f read through f(fine DOF):
// the 27 fine nodes 2N + a in the order (x, y, z), weight 1 on the node itself and 1/2 per axis
// beside it. A fine node outside the grid is read at the clamped position with weight 0, so
// the 81 loads carry no branch and issue together.
template <typename F>
__device__ __forceinline__ void mg_restrict_sum(F f, const MgGrid& fg, int I, int J, int K,
                                                float (&acc)[3]) {
  acc[0] = acc[1] = acc[2] = 0.f;
#pragma unroll
  for (int ax = -1; ax <= 1; ++ax)
#pragma unroll
    for (int ay = -1; ay <= 1; ++ay)
#pragma unroll
      for (int az = -1; az <= 1; ++az) {
        const int mx = 2 * I + ax, my = 2 * J + ay, mz = 2 * K + az;
        const bool ok = mx >= 0 && mx < fg.nnx && my >= 0 && my < fg.nny && mz >= 0 && mz < fg.nnz;
        const float w = ok ? (ax ? 0.5f : 1.f) * (ay ? 0.5f : 1.f) * (az ? 0.5f : 1.f) : 0.f;
        const int64_t m = mg_node(fg, min(max(mx, 0), fg.nnx - 1), min(max(my, 0), fg.nny - 1),
                                  min(max(mz, 0), fg.nnz - 1));
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] = fmaf(w, f(3 * m + c), acc[c]);
      }
}

// Prolongation sum (P x_c)(n) at fine node n = (i, j, k), x_c read through f(coarse DOF): the
// (up to) eight trilinear parents in the order (x, y, z); an even coordinate has one parent,
// and its second parent repeats the first with weight 0.
template <typename F>
__device__ __forceinline__ void mg_prolong_sum(F f, const MgGrid& cg, int i, int j, int k,
                                               float (&acc)[3]) {
  const int i0 = i >> 1, j0 = j >> 1, k0 = k >> 1;
  const int i1 = (i & 1) ? i0 + 1 : i0, j1 = (j & 1) ? j0 + 1 : j0, k1 = (k & 1) ? k0 + 1 : k0;
  const float wi[2] = {(i & 1) ? 0.5f : 1.f, (i & 1) ? 0.5f : 0.f};
  const float wj[2] = {(j & 1) ? 0.5f : 1.f, (j & 1) ? 0.5f : 0.f};
  const float wk[2] = {(k & 1) ? 0.5f : 1.f, (k & 1) ? 0.5f : 0.f};
  const int ci[2] = {i0, i1}, cj[2] = {j0, j1}, ck[2] = {k0, k1};
  acc[0] = acc[1] = acc[2] = 0.f;
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int64_t N = mg_node(cg, ci[a], cj[b], ck[c]);
        const float w = wi[a] * wj[b] * wk[c];
#pragma unroll
        for (int q = 0; q < 3; ++q) acc[q] = fmaf(w, f(3 * N + q), acc[q]);
      }
}

// Restriction r_c = P^T r_f at coarse node N; with wdc, also the coarse level's
// pre-smoothing step from zero, x0c = wdc * r_c.
__device__ __forceinline__ void mg_restrict_node(const float* __restrict__ rf, const MgGrid& fg,
                                                 float* __restrict__ rc, const MgGrid& cg,
                                                 int64_t N, const float* __restrict__ wdc,
                                                 float* __restrict__ x0c) {
  int I, J, K;
  mg_decode(N, cg, I, J, K);
  float acc[3];
  mg_restrict_sum([&](int64_t q) { return rf[q]; }, fg, I, J, K, acc);
#pragma unroll
  for (int c = 0; c < 3; ++c) rc[3 * N + c] = acc[c];
  if (wdc) {
#pragma unroll
    for (int c = 0; c < 3; ++c) x0c[3 * N + c] = wdc[3 * N + c] * acc[c];
  }
}

// Prolongation x_f += P x_c at fine node n; DOFs with fixcnt > 0 (level 0 only; null on the
// coarse levels) get no correction.
__device__ __forceinline__ void mg_prolong_node(const float* __restrict__ xc, const MgGrid& cg,
                                                float* __restrict__ xf, const MgGrid& fg,
                                                int64_t n, const uint8_t* __restrict__ fixcnt) {
  int i, j, k;
  mg_decode(n, fg, i, j, k);
  float acc[3];
  mg_prolong_sum([&](int64_t q) { return xc[q]; }, cg, i, j, k, acc);
#pragma unroll
  for (int q = 0; q < 3; ++q)
    if (!fixcnt || !fixcnt[3 * n + q]) xf[3 * n + q] += acc[q];
}

// Dense rows x[row0 + u*rs] = Ainv[row, :] r, u < 8, for the calling wave: lanes split the
// columns and the eight rows' loads are in flight together.
__device__ __forceinline__ void mg_dense_rows8(const float* __restrict__ Ainv, int n,
                                               const float* __restrict__ r,
                                               float* __restrict__ x, int row0, int rs) {
  const int lane = threadIdx.x & 63;
  float acc[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) acc[u] = 0.f;
  for (int c = lane; c < n; c += 64) {
    const float rc = r[c];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int row = min(row0 + u * rs, n - 1);  // rows past n repeat the last, unused
      acc[u] = fmaf(Ainv[(int64_t)row * n + c], rc, acc[u]);
    }
  }
#pragma unroll
  for (int u = 0; u < 8; ++u) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc[u] += __shfl_xor(acc[u], o, 64);
  }
#pragma unroll
  for (int u = 0; u < 8; ++u)
    if (lane == 0 && row0 + u * rs < n) x[row0 + u * rs] = acc[u];
}

// Stencil-level block mapping: a block handles kBlock consecutive nodes for one row component
// r (uniform in the block, so the coefficient offsets stay scalar). Hardware block b runs on
// XCD b % 8; the (node range, r) pairs are dealt out so that XCD x gets a contiguous run of
// node ranges with all three rows of each, and the forward blocks, the neighbours' transposed
// blocks and the three row passes over them come through that XCD's L2 (numbering the pairs
// (range, r) = (b mod R, b div R) instead fetched each stencil block from HBM ~3 times).
// Grid: 8 * per blocks with per = ceil(3 * ceil(n_nodes / kBlock) / 8).
__device__ __forceinline__ bool mg_row_block(const MgGrid& g, int64_t& n, int& r) {
  const int nr = (int)((g.n_nodes + kBlock - 1) / kBlock);
  const int per = (3 * nr + 7) / 8;
  const int logical = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  if (logical >= 3 * nr) return false;
  r = logical % 3;
  n = (int64_t)(logical / 3) * kBlock + threadIdx.x;
  return n < g.n_nodes;
}
inline unsigned mg_row_grid(int64_t n_nodes) {
  const int64_t nr = (n_nodes + kBlock - 1) / kBlock;
  return 8u * (unsigned)((3 * nr + 7) / 8);
}

// Plane-split stencil sweeps: a block takes 64 consecutive nodes of one row component r, and
// wave w the neighbour plane di = w - 1: its 27 coefficients and 27 neighbour values, 54 loads
// that leave in one batch (a whole row per thread is 162 loads in three). Wave 0 adds the
// three plane partials in fixed order. (node chunk, r) pairs are dealt XCD-contiguously.
// Grid: 8 * ceil(3 * ceil(n_nodes / 64) / 8) blocks of 192 threads.
constexpr int kPlaneNodes = 64;
__device__ __forceinline__ bool mg_plane_block(const MgGrid& g, int& n, int& r) {
  const int nr = (int)((g.n_nodes + kPlaneNodes - 1) / kPlaneNodes);
  const int per = (3 * nr + 7) / 8;
  const int logical = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  if (logical >= 3 * nr) return false;
  r = logical % 3;
  n = (logical / 3) * kPlaneNodes + (int)(threadIdx.x & 63);
  return true;
}
inline unsigned mg_plane_grid(int64_t n_nodes) {
  const int64_t nr = (n_nodes + kPlaneNodes - 1) / kPlaneNodes;
  return 8u * (unsigned)((3 * nr + 7) / 8);
}

// Partial of row r at node n = (i, j, k) over the neighbour plane DI: the sum over (dj, dk, c)
// of A(n, n + (DI, dj, dk))[r][c] x(n + (DI, dj, dk))[c], with x zero outside the grid (read
// at n there, like a backward block's coefficients).
template <int DI>
__device__ __forceinline__ float mg_plane_row(const float* __restrict__ A, const MgGrid& g, int n,
                                              int i, int j, int k, int r,
                                              const float* __restrict__ x) {
  const int sj = g.nnz, si = g.nny * g.nnz;
  const int64_t nn = g.n_nodes;
  float p = 0.f;
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    const int dj = e / 3 - 1, dk = e % 3 - 1;
    const int d = (DI + 1) * 9 + (dj + 1) * 3 + (dk + 1);
    const bool ok = (i + DI) >= 0 && (i + DI) < g.nnx && (j + dj) >= 0 && (j + dj) < g.nny &&
                    (k + dk) >= 0 && (k + dk) < g.nnz;
    const int m = ok ? n + DI * si + dj * sj + dk : n;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float xv = x[(unsigned)(3 * m + c)];
      const float co = d >= 13 ? A[(int64_t)((d - 13) * 9 + 3 * r + c) * nn + n]
                               : A[(int64_t)((13 - d) * 9 + 3 * c + r) * nn + m];
      p = fmaf(co, ok ? xv : 0.f, p);
    }
  }
  return p;
}
// y = (A x)[3n + r] for the calling block's chunk, from the three waves; the returned value
// is valid in wave 0 only.
__device__ __forceinline__ float mg_plane_apply(const float* __restrict__ A, const MgGrid& g,
                                                int n, int r, const float* __restrict__ x,
                                                float (*part)[kPlaneNodes]) {
  const int w = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
  const int nc = n < g.n_nodes ? n : 0;  // past the grid: node 0 (in bounds, unused)
  const int k = nc % g.nnz, t = nc / g.nnz, j = t % g.nny, i = t / g.nny;
  float p;
  if (w == 0)
    p = mg_plane_row<-1>(A, g, nc, i, j, k, r, x);
  else if (w == 1)
    p = mg_plane_row<0>(A, g, nc, i, j, k, r, x);
  else
    p = mg_plane_row<1>(A, g, nc, i, j, k, r, x);
  if (w > 0) part[w - 1][lane] = p;
  __syncthreads();
  return (p + part[0][lane]) + part[1][lane];
}

__global__ __launch_bounds__(3 * kPlaneNodes) void k_mg_pre3(const float* __restrict__ A,
                                                            MgGrid g,
                                                            const float* __restrict__ rhs,
                                                            const float* __restrict__ x0,
                                                            float* __restrict__ t) {
  __shared__ float part[2][kPlaneNodes];
  int n, r;
  if (!mg_plane_block(g, n, r)) return;  // whole block: no barrier follows for it
  const float y = mg_plane_apply(A, g, n, r, x0, part);
  if (threadIdx.x < kPlaneNodes && n < g.n_nodes) {
    const unsigned q = 3 * n + r;
    t[q] = rhs[q] - y;
  }
}

__global__ __launch_bounds__(3 * kPlaneNodes) void k_mg_post3(const float* __restrict__ A,
                                                             MgGrid g,
                                                             const float* __restrict__ wd,
                                                             const float* __restrict__ rhs,
                                                             const float* __restrict__ x,
                                                             float* __restrict__ out) {
  __shared__ float part[2][kPlaneNodes];
  int n, r;
  if (!mg_plane_block(g, n, r)) return;  // whole block: no barrier follows for it
  const float y = mg_plane_apply(A, g, n, r, x, part);
  if (threadIdx.x < kPlaneNodes && n < g.n_nodes) {
    const unsigned q = 3 * n + r;
    out[q] = x[q] + wd[q] * (rhs[q] - y);
  }
}

// One step of the Chebyshev smoother on a stencil level: r = mi (rhs - A x), d' = a d + b r
// (d ignored when a = 0), x' = x + d'. d_out may be null (last step).
__global__ __launch_bounds__(3 * kPlaneNodes) void k_mg_cheb(
    const float* __restrict__ A, MgGrid g, const float* __restrict__ mi,
    const float* __restrict__ rhs, const float* __restrict__ x, const float* __restrict__ d,
    float a, float b, float* __restrict__ x_out, float* __restrict__ d_out) {
  __shared__ float part[2][kPlaneNodes];
  int n, r;
  if (!mg_plane_block(g, n, r)) return;  // whole block: no barrier follows for it
  const float y = mg_plane_apply(A, g, n, r, x, part);
  if (threadIdx.x < kPlaneNodes && n < g.n_nodes) {
    const unsigned q = 3 * n + r;
    const float rr = mi[q] * (rhs[q] - y);
    const float dn = a != 0.f ? fmaf(a, d[q], b * rr) : b * rr;
    x_out[q] = x[q] + dn;
    if (d_out) d_out[q] = dn;
  }
}

__global__ __launch_bounds__(kBlock) void k_mg_pre(const float* __restrict__ A, MgGrid g,
                                                   const float* __restrict__ rhs,
                                                   const float* __restrict__ x0,
                                                   float* __restrict__ t) {
  int64_t n;
  int r;
  if (!mg_row_block(g, n, r)) return;
  mg_pre_row(A, g, rhs, x0, t, n, r);
}

// k_mg_pre / k_mg_post on the half-precision blocks (k_mg_stencil_f16).
__global__ __launch_bounds__(kBlock) void k_mg_pre_h(const _Float16* __restrict__ Ah,
                                                     const float* __restrict__ S, MgGrid g,
                                                     const float* __restrict__ rhs,
                                                     const float* __restrict__ x0,
                                                     float* __restrict__ t) {
  int64_t n;
  int r;
  if (!mg_row_block(g, n, r)) return;
  int i, j, k;
  mg_decode(n, g, i, j, k);
  float nb[27][3];
  mg_gather(g, n, i, j, k, [&](int64_t q) { return x0[q]; }, nb);
  const float y = stencil_row_h(Ah, S, g, n, i, j, k, r, nb);
  const int64_t q = 3 * n + r;
  t[q] = rhs[q] - y;
}

__global__ __launch_bounds__(kBlock) void k_mg_post_h(const _Float16* __restrict__ Ah,
                                                      const float* __restrict__ S, MgGrid g,
                                                      const float* __restrict__ wd,
                                                      const float* __restrict__ rhs,
                                                      const float* __restrict__ x,
                                                      float* __restrict__ out) {
  int64_t n;
  int r;
  if (!mg_row_block(g, n, r)) return;
  int i, j, k;
  mg_decode(n, g, i, j, k);
  float nb[27][3];
  mg_gather(g, n, i, j, k, [&](int64_t q) { return x[q]; }, nb);
  const float y = stencil_row_h(Ah, S, g, n, i, j, k, r, nb);
  const int64_t q = 3 * n + r;
  out[q] = x[q] + wd[q] * (rhs[q] - y);
}

__global__ __launch_bounds__(kBlock) void k_mg_post(const float* __restrict__ A, MgGrid g,
                                                    const float* __restrict__ wd,
                                                    const float* __restrict__ rhs,
                                                    const float* __restrict__ x,
                                                    float* __restrict__ out) {
  int64_t n;
  int r;
  if (!mg_row_block(g, n, r)) return;
  mg_post_row(A, g, wd, rhs, x, out, n, r);
}

// Restriction (coarse node per thread).
__global__ __launch_bounds__(kBlock) void k_mg_restrict(const float* __restrict__ rf,
                                                        MgGrid fg, float* __restrict__ rc,
                                                        MgGrid cg, const float* __restrict__ wdc,
                                                        float* __restrict__ x0c) {
  const int64_t N = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (N >= cg.n_nodes) return;
  mg_restrict_node(rf, fg, rc, cg, N, wdc, x0c);
}

// Prolongation (fine node per thread).
__global__ __launch_bounds__(kBlock) void k_mg_prolong(const float* __restrict__ xc, MgGrid cg,
                                                       float* __restrict__ xf, MgGrid fg,
                                                       const uint8_t* __restrict__ fixcnt) {
  const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (n >= fg.n_nodes) return;
  mg_prolong_node(xc, cg, xf, fg, n, fixcnt);
}

// Coarsest level: x = Ainv r (dense, n <= 192); eight rows per wave, four waves per block.
__global__ __launch_bounds__(kBlock) void k_mg_dense(const float* __restrict__ Ainv, int n,
                                                     const float* __restrict__ r,
                                                     float* __restrict__ x) {
  const int row0 = blockIdx.x * 8 * kWaves + (threadIdx.x >> 6);
  if (row0 >= n) return;
  mg_dense_rows8(Ainv, n, r, x, row0, kWaves);
}

// ---------------------------------------------------------------------------------------
// Fused coarse-level V-cycle. On the small stencil levels a launch costs its dispatch (~1.7 us
// in a graph) plus its dependent memory round trips, more than its arithmetic: a one-row-per-
// thread stencil sweep issues 162 loads per thread, which leave in three batches of at most 63
// outstanding loads. From level f on, each level's work is therefore two launches over bricks
// of at most 4x4x4 nodes, and the dense level one more:
//   k_mg_down   (level l): rhs_l = P^T t_{l-1} on the brick and its one-node shell, x0 = wd rhs_l
//                (the pre-smoothing step from zero), t_l = rhs_l - A_l x0 on the brick;
//   k_mg_rdense (dense level L-1): x_{L-1} = Ainv P^T t_{L-2};
//   k_mg_up     (level l): x = wd rhs_l + P c_{l+1} on the brick and its shell (c: the coarse
//                correction), post-smoothing t_l = x + wd (rhs_l - A_l x) on the brick.
// A block has 576 threads, one per (row component r, plane di, brick node): the thread's 27
// stencil coefficients (the 3x3 neighbour blocks at offset di) are loaded first and arrive
// together with the brick's input vectors, so a launch waits for one memory round trip. The
// restriction is separable and runs in LDS on the fine region behind the shell, whose values
// are each loaded once. The three plane partials of a row are added in LDS in fixed order.
// The arithmetic is the same linear map as the one-step kernels above, summed in another order.
// ---------------------------------------------------------------------------------------
constexpr int kFB = 4;                   // brick extent per axis (coarse-level nodes)
constexpr int kFBN = kFB * kFB * kFB;    // 64 brick nodes
constexpr int kFS = kFB + 2;             // brick plus shell, per axis
constexpr int kFSN = kFS * kFS * kFS;    // 216
constexpr int kFF = 2 * kFB + 5;         // fine region behind the shell, per axis
constexpr int kFusedThreads = 9 * kFBN;  // 576
constexpr int kFuseMaxNodes = 16384;     // levels fused (at most this many nodes)

// Bricks per axis; the bricks tile the level from node 0, the last ones reach past the grid.
struct MgBrick {
  int nbi, nbj, nbk;
};

// Chebyshev mode of the fused kernels (mi null: damped Jacobi). The brick rows take one
// smoother step r = mi (rhs - A x), d' = a d + b r, x' = x + d' and store x' and d'; k_mg_up
// starts from x = xpre + P c (xpre: the level's pre-smoothed iterate) with a = 0.
struct MgCheb {
  const float* mi = nullptr;
  float a = 0.f, b = 0.f;
  float* x_out = nullptr;
  float* d_out = nullptr;
  const float* xpre = nullptr;
};

// Brick of this block, dealt XCD-contiguously (hardware block b runs on XCD b % 8; XCD x takes
// a contiguous run of bricks, so neighbouring bricks' shells are read through one L2).
// Grid: 8 * ceil(n_bricks / 8) blocks.
__device__ __forceinline__ bool mg_brick(const MgBrick& b, int& i0, int& j0, int& k0) {
  const int nb = b.nbi * b.nbj * b.nbk;
  const int per = (nb + 7) / 8;
  const int logical = (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
  if (logical >= nb) return false;
  k0 = (logical % b.nbk) * kFB;
  const int t = logical / b.nbk;
  j0 = (t % b.nbj) * kFB;
  i0 = (t / b.nbj) * kFB;
  return true;
}

// This thread's plane task: row component r and neighbour plane di (both uniform per wave)
// of brick node (li, lj, lk); valid when the node lies inside the grid.
struct MgTask {
  int r, di, li, lj, lk;
  bool valid;
};
__device__ __forceinline__ MgTask mg_task(const MgGrid& g, int i0, int j0, int k0) {
  MgTask t;
  const int rd = (int)threadIdx.x / kFBN, ln = (int)threadIdx.x % kFBN;
  t.r = rd / 3;
  t.di = rd % 3 - 1;
  t.lk = ln % kFB;
  t.lj = (ln / kFB) % kFB;
  t.li = ln / (kFB * kFB);
  t.valid = i0 + t.li < g.nnx && j0 + t.lj < g.nny && k0 + t.lk < g.nnz;
  return t;
}

// Coefficients A(n, n + (DI, dj, dk))[r][c] for the nine (dj, dk), c fastest; a backward block
// whose neighbour is outside the grid is read at n (finite; its x is zero). 32-bit offsets: the
// fused levels hold at most kFuseMaxNodes nodes.
template <int DI>
__device__ __forceinline__ void mg_plane_coefs(const float* __restrict__ A, const MgGrid& g,
                                               int n, int i, int j, int k, int r,
                                               float (&co)[27]) {
  const int sj = g.nnz, si = g.nny * g.nnz, nn = (int)g.n_nodes;
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    const int dj = e / 3 - 1, dk = e % 3 - 1;
    const int d = (DI + 1) * 9 + (dj + 1) * 3 + (dk + 1);
    if (d >= 13) {
#pragma unroll
      for (int c = 0; c < 3; ++c) co[3 * e + c] = A[(unsigned)(((d - 13) * 9 + 3 * r + c) * nn + n)];
    } else {
      const bool ok = (i + DI) >= 0 && (i + DI) < g.nnx && (j + dj) >= 0 && (j + dj) < g.nny &&
                      (k + dk) >= 0 && (k + dk) < g.nnz;
      const int m = ok ? n + DI * si + dj * sj + dk : n;
#pragma unroll
      for (int c = 0; c < 3; ++c) co[3 * e + c] = A[(unsigned)(((13 - d) * 9 + 3 * c + r) * nn + m)];
    }
  }
}
__device__ __forceinline__ void mg_task_coefs(const float* __restrict__ A, const MgGrid& g,
                                              const MgTask& t, int i0, int j0, int k0,
                                              float (&co)[27]) {
  // an invalid task loads node 0's coefficients (in bounds, never used)
  const int i = t.valid ? i0 + t.li : 0, j = t.valid ? j0 + t.lj : 0, k = t.valid ? k0 + t.lk : 0;
  const int n = (i * g.nny + j) * g.nnz + k;
  if (t.di < 0)
    mg_plane_coefs<-1>(A, g, n, i, j, k, t.r, co);
  else if (t.di == 0)
    mg_plane_coefs<0>(A, g, n, i, j, k, t.r, co);
  else
    mg_plane_coefs<1>(A, g, n, i, j, k, t.r, co);
}

// The three components of node m (32-bit node index) of a node-major vector.
struct F3 {
  float v[3];
};
__device__ __forceinline__ F3 mg_load3(const float* __restrict__ x, int m) {
  const float* p = x + (unsigned)(3 * m);
  F3 r;
  r.v[0] = p[0];
  r.v[1] = p[1];
  r.v[2] = p[2];
  return r;
}

// Plane partial of the task's row: sum over (dj, dk, c) of co * x, with x read from the shell
// array xs (node-major, 3 floats a node, zeros outside the grid).
__device__ __forceinline__ float mg_task_partial(const float (&co)[27], const float* xs,
                                                 const MgTask& t) {
  const int q0 = ((t.li + 1 + t.di) * kFS + (t.lj + 1)) * kFS + (t.lk + 1);
  float p = 0.f;
#pragma unroll
  for (int e = 0; e < 9; ++e) {
    const int q = q0 + (e / 3 - 1) * kFS + (e % 3 - 1);
#pragma unroll
    for (int c = 0; c < 3; ++c) p = fmaf(co[3 * e + c], xs[3 * q + c], p);
  }
  return p;
}

// Shell position of brick node (li, lj, lk).
__device__ __forceinline__ int mg_shell_q(int li, int lj, int lk) {
  return ((li + 1) * kFS + (lj + 1)) * kFS + (lk + 1);
}

// The brick's rows from the plane partials P (index (3r + di + 1) * 64 + brick node): threads
// below 192 take row (brick node, r = threadIdx / 64) and call f(n, r, shell position, y).
template <typename Fn>
__device__ __forceinline__ void mg_brick_rows(const float* P, const MgGrid& g, int i0, int j0,
                                              int k0, Fn f) {
  if (threadIdx.x >= 3 * kFBN) return;
  const int r = (int)threadIdx.x / kFBN, ln = (int)threadIdx.x % kFBN;
  const int lk = ln % kFB, lj = (ln / kFB) % kFB, li = ln / (kFB * kFB);
  if (i0 + li >= g.nnx || j0 + lj >= g.nny || k0 + lk >= g.nnz) return;
  const float y = (P[(3 * r) * kFBN + ln] + P[(3 * r + 1) * kFBN + ln]) + P[(3 * r + 2) * kFBN + ln];
  f(((i0 + li) * g.nny + j0 + lj) * g.nnz + k0 + lk, r, mg_shell_q(li, lj, lk), y);
}

__global__ __launch_bounds__(kFusedThreads) void k_mg_down(
    const float* __restrict__ tf, MgGrid fg, const float* __restrict__ A,
    const float* __restrict__ wd, MgGrid g, MgBrick b, float* __restrict__ rhs,
    float* __restrict__ t_out, MgCheb ch) {
  // fine region F (x, y, z), z-restricted G (x, y, kz), y-restricted H (x, jy, kz) in F's
  // space, shell rhs R and x0 X, plane partials P; 3 floats a node
  __shared__ float F[3 * kFF * kFF * kFF], G[3 * kFF * kFF * kFS];
  __shared__ float R[3 * kFSN], X[3 * kFSN], P[9 * kFBN];
  int i0, j0, k0;
  if (!mg_brick(b, i0, j0, k0)) return;  // whole block: no barrier follows for it
  const MgTask tk = mg_task(g, i0, j0, k0);
  float co[27];
  mg_task_coefs(A, g, tk, i0, j0, k0, co);
  // the fine region (zeros outside the fine grid; one node per thread and step) and the shell's
  // smoother weights (zeros outside the grid), issued behind the coefficients
  const int fx0 = 2 * i0 - 3, fy0 = 2 * j0 - 3, fz0 = 2 * k0 - 3;
  constexpr int kNF = kFF * kFF * kFF;
  constexpr int kFLoads = (kNF + kFusedThreads - 1) / kFusedThreads;
  F3 fv[kFLoads];
#pragma unroll
  for (int u = 0; u < kFLoads; ++u) {
    const int q = threadIdx.x + u * kFusedThreads;
    const int x = fx0 + q / (kFF * kFF), y = fy0 + (q / kFF) % kFF, z = fz0 + q % kFF;
    fv[u].v[0] = fv[u].v[1] = fv[u].v[2] = 0.f;
    if (q < kNF && x >= 0 && x < fg.nnx && y >= 0 && y < fg.nny && z >= 0 && z < fg.nnz)
      fv[u] = mg_load3(tf, (x * fg.nny + y) * fg.nnz + z);
  }
  F3 wv;
  wv.v[0] = wv.v[1] = wv.v[2] = 0.f;
  {
    const int q = threadIdx.x;
    const int I = i0 - 1 + q / (kFS * kFS), J = j0 - 1 + (q / kFS) % kFS, K = k0 - 1 + q % kFS;
    if (q < kFSN && I >= 0 && I < g.nnx && J >= 0 && J < g.nny && K >= 0 && K < g.nnz)
      wv = mg_load3(wd, (I * g.nny + J) * g.nnz + K);
  }
#pragma unroll
  for (int u = 0; u < kFLoads; ++u) {
    const int q = threadIdx.x + u * kFusedThreads;
    if (q < kNF) {
#pragma unroll
      for (int c = 0; c < 3; ++c) F[3 * q + c] = fv[u].v[c];
    }
  }
  __syncthreads();
  // separable restriction (weights 1/2, 1, 1/2): z, then y, then x; a pass task is (outer
  // index, o, component), the steps of a thread unrolled so their LDS reads overlap
  constexpr int kZ = 3 * kFF * kFF * kFS, kY = 3 * kFF * kFS * kFS;
#pragma unroll
  for (int u = 0; u < (kZ + kFusedThreads - 1) / kFusedThreads; ++u) {
    const int q = threadIdx.x + u * kFusedThreads;
    if (q < kZ) {
      const int c = q % 3, v = q / 3, o = v % kFS, xy = v / kFS;
      const int ib = 3 * (xy * kFF + 2 * o) + c;
      G[q] = fmaf(0.5f, F[ib + 6], 0.5f * F[ib] + F[ib + 3]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < (kY + kFusedThreads - 1) / kFusedThreads; ++u) {
    const int q = threadIdx.x + u * kFusedThreads;
    if (q < kY) {
      const int c = q % 3, v = q / 3, kz = v % kFS, o = (v / kFS) % kFS, x = v / (kFS * kFS);
      const int ib = 3 * ((x * kFF + 2 * o) * kFS + kz) + c;
      F[q] = fmaf(0.5f, G[ib + 6 * kFS], 0.5f * G[ib] + G[ib + 3 * kFS]);
    }
  }
  __syncthreads();
  // x: one shell node per thread, then x0 = wd rhs there; the brick's rhs is stored
  if (threadIdx.x < kFSN) {
    const int q = threadIdx.x;
    const int li = q / (kFS * kFS), lj = (q / kFS) % kFS, lk = q % kFS;
    const int I = i0 - 1 + li, J = j0 - 1 + lj, K = k0 - 1 + lk;
    const bool own = li >= 1 && li <= kFB && lj >= 1 && lj <= kFB && lk >= 1 && lk <= kFB &&
                     I < g.nnx && J < g.nny && K < g.nnz;
    float* dst = rhs + (unsigned)(3 * ((I * g.nny + J) * g.nnz + K));
    const int ib = 3 * (2 * li * kFS * kFS + lj * kFS + lk);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const float rv =
          fmaf(0.5f, F[ib + 6 * kFS * kFS + c], 0.5f * F[ib + c] + F[ib + 3 * kFS * kFS + c]);
      R[3 * q + c] = rv;
      X[3 * q + c] = wv.v[c] * rv;
      if (own) dst[c] = rv;
    }
  }
  __syncthreads();
  P[threadIdx.x] = tk.valid ? mg_task_partial(co, X, tk) : 0.f;
  __syncthreads();
  mg_brick_rows(P, g, i0, j0, k0, [&](int n, int r, int q, float y) {
    const unsigned e = 3 * n + r;
    if (ch.mi) {  // second Chebyshev step (X holds x1 = d1)
      const float x1 = X[3 * q + r];
      const float dn = fmaf(ch.a, x1, ch.b * (ch.mi[e] * (R[3 * q + r] - y)));
      ch.x_out[e] = x1 + dn;
      ch.d_out[e] = dn;
    } else {
      t_out[e] = R[3 * q + r] - y;
    }
  });
}

__global__ __launch_bounds__(kFusedThreads) void k_mg_up(
    const float* __restrict__ xc, MgGrid cg, const float* __restrict__ A,
    const float* __restrict__ wd, const float* __restrict__ rhs, MgGrid g, MgBrick b,
    float* __restrict__ out, MgCheb ch) {
  // coarse correction region C (the parents of the shell), shell rhs, wd and x, partials P
  constexpr int kC = kFB / 2 + 3;
  __shared__ float C[3 * kC * kC * kC], RS[3 * kFSN], WS[3 * kFSN], X[3 * kFSN], P[9 * kFBN];
  int i0, j0, k0;
  if (!mg_brick(b, i0, j0, k0)) return;  // whole block: no barrier follows for it
  // coarse parents of the in-grid shell nodes lie in [max(i0 - 1, 0) / 2, that + kC)
  const int ci0 = max(i0 - 1, 0) >> 1, cj0 = max(j0 - 1, 0) >> 1, ck0 = max(k0 - 1, 0) >> 1;
  const MgTask tk = mg_task(g, i0, j0, k0);
  float co[27];
  mg_task_coefs(A, g, tk, i0, j0, k0, co);
  F3 cv, rv, wv;
#pragma unroll
  for (int c = 0; c < 3; ++c) cv.v[c] = rv.v[c] = wv.v[c] = 0.f;
  if (threadIdx.x < kC * kC * kC) {
    const int q = threadIdx.x;
    const int x = ci0 + q / (kC * kC), y = cj0 + (q / kC) % kC, z = ck0 + q % kC;
    if (x < cg.nnx && y < cg.nny && z < cg.nnz) cv = mg_load3(xc, (x * cg.nny + y) * cg.nnz + z);
  }
  const int q = threadIdx.x;
  const int I = i0 - 1 + q / (kFS * kFS), J = j0 - 1 + (q / kFS) % kFS, K = k0 - 1 + q % kFS;
  const bool in = q < kFSN && I >= 0 && I < g.nnx && J >= 0 && J < g.nny && K >= 0 && K < g.nnz;
  if (in) {
    const int m = (I * g.nny + J) * g.nnz + K;
    rv = mg_load3(rhs, m);
    wv = mg_load3(ch.mi ? ch.xpre : wd, m);  // Chebyshev: the pre-smoothed x itself
  }
  if (threadIdx.x < kC * kC * kC) {
#pragma unroll
    for (int c = 0; c < 3; ++c) C[3 * threadIdx.x + c] = cv.v[c];
  }
  __syncthreads();
  // x = wd rhs + P c on the shell (zeros outside the grid)
  if (q < kFSN) {
    float acc[3] = {0.f, 0.f, 0.f};
    if (in) {
      const int pa[2] = {(I >> 1) - ci0, ((I + 1) >> 1) - ci0};
      const int pb[2] = {(J >> 1) - cj0, ((J + 1) >> 1) - cj0};
      const int pc[2] = {(K >> 1) - ck0, ((K + 1) >> 1) - ck0};
      const float fa[2] = {(I & 1) ? 0.5f : 1.f, (I & 1) ? 0.5f : 0.f};
      const float fb[2] = {(J & 1) ? 0.5f : 1.f, (J & 1) ? 0.5f : 0.f};
      const float fc[2] = {(K & 1) ? 0.5f : 1.f, (K & 1) ? 0.5f : 0.f};
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
          for (int w = 0; w < 2; ++w) {
            const float wt = fa[s] * fb[t] * fc[w];
            const int e = 3 * ((pa[s] * kC + pb[t]) * kC + pc[w]);
#pragma unroll
            for (int c = 0; c < 3; ++c) acc[c] = fmaf(wt, C[e + c], acc[c]);
          }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      X[3 * q + c] = in ? (ch.mi ? wv.v[c] : wv.v[c] * rv.v[c]) + acc[c] : 0.f;
      RS[3 * q + c] = rv.v[c];
      WS[3 * q + c] = wv.v[c];
    }
  }
  __syncthreads();
  P[threadIdx.x] = tk.valid ? mg_task_partial(co, X, tk) : 0.f;
  __syncthreads();
  mg_brick_rows(P, g, i0, j0, k0, [&](int n, int r, int q, float y) {
    const int e = 3 * q + r;
    const unsigned o = 3 * n + r;
    if (ch.mi) {  // first post-smoothing Chebyshev step
      const float dn = ch.b * (ch.mi[o] * (RS[e] - y));
      ch.x_out[o] = X[e] + dn;
      ch.d_out[o] = dn;
    } else {
      out[o] = X[e] + WS[e] * (RS[e] - y);
    }
  });
}

// Dense level: x = Ainv P^T t. Block b takes rows [16b, 16b + 16): its threads below nc form
// the whole right-hand side (27 fine values each), each of its four waves four rows (lanes over
// columns lane + 64v, then the butterfly sum). Grid: ceil(nc / 16) blocks of 256 threads.
constexpr int kDenseMax = 192;
__global__ __launch_bounds__(kBlock) void k_mg_rdense(const float* __restrict__ tf, MgGrid fg,
                                                      MgGrid cg, const float* __restrict__ Ainv,
                                                      int nc, float* __restrict__ x) {
  __shared__ float rc[kDenseMax];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row0 = blockIdx.x * 16 + 4 * w;
  float av[4][3];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const int row = row0 + u, col = lane + 64 * v;
      av[u][v] = (row < nc && col < nc) ? Ainv[row * nc + col] : 0.f;
    }
  if ((int)threadIdx.x < nc) {
    const int c = threadIdx.x % 3, N = threadIdx.x / 3;
    const int K = N % cg.nnz, J = (N / cg.nnz) % cg.nny, I = N / (cg.nnz * cg.nny);
    float acc = 0.f;
#pragma unroll
    for (int ax = -1; ax <= 1; ++ax)
#pragma unroll
      for (int ay = -1; ay <= 1; ++ay)
#pragma unroll
        for (int az = -1; az <= 1; ++az) {
          const int mx = 2 * I + ax, my = 2 * J + ay, mz = 2 * K + az;
          const bool ok = mx >= 0 && mx < fg.nnx && my >= 0 && my < fg.nny && mz >= 0 && mz < fg.nnz;
          const float wt = ok ? (ax ? 0.5f : 1.f) * (ay ? 0.5f : 1.f) * (az ? 0.5f : 1.f) : 0.f;
          const int m = (min(max(mx, 0), fg.nnx - 1) * fg.nny + min(max(my, 0), fg.nny - 1)) * fg.nnz +
                        min(max(mz, 0), fg.nnz - 1);
          acc = fmaf(wt, tf[3 * m + c], acc);
        }
    rc[threadIdx.x] = acc;
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    float acc = 0.f;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const int col = lane + 64 * v;
      if (col < nc) acc = fmaf(av[u][v], rc[col], acc);
    }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0 && row0 + u < nc) x[row0 + u] = acc;
  }
}

// Dense inverse of the coarsest Galerkin operator, on the device in one 1024-thread block
// (FP64), written as FP32 row-major n x n. The matrix is assembled from the level's symmetric
// half stencil (neighbour d >= 13 from the node's own blocks, d < 13 as the transpose of the
// neighbour's block 13 - d) and inverted by in-place Gauss-Jordan elimination (the sweep
// operator): for pivot k with p = a_kk, a_ij -= a_ik a_kj / p, a_kj = a_kj / p,
// a_ik = -a_ik / p, a_kk = 1 / p; after all n pivots the array holds the inverse. No pivoting:
// the operator is symmetric positive semidefinite, and its pivots are the Schur-complement
// diagonals of the Cholesky factorisation. A pivot that rounding drives to <= 1e-13 of its
// original diagonal (a DOF no element stiffens, or a floating piece) decouples its DOF: row and
// column k are zeroed and the DOF gets the Jacobi value 1 / a_kk (1 if a_kk <= 0), so the
// result stays symmetric positive definite.
//
// Thread (tx, ty) of the 32 x 32 grid holds entries (ty + 32 a, tx + 32 b), a, b < R, in
// registers; per pivot it reads R values of the pivot column and R of the pivot row from LDS,
// so a step costs 2R LDS reads and R^2 FMAs per thread and one barrier. n <= 32 R; the padded
// rows and columns hold the identity and never couple.
__device__ __forceinline__ double mg_dense_entry(const float* __restrict__ A, const MgGrid& g,
                                                 int r, int c) {
  const int64_t N = r / 3, M = c / 3;
  int I, J, K, MI, MJ, MK;
  mg_decode(N, g, I, J, K);
  mg_decode(M, g, MI, MJ, MK);
  const int di = MI - I, dj = MJ - J, dk = MK - K;
  if (di < -1 || di > 1 || dj < -1 || dj > 1 || dk < -1 || dk > 1) return 0.0;
  const int d = (di + 1) * 9 + (dj + 1) * 3 + (dk + 1);
  const int ci = r % 3, cj = c % 3;
  return d >= 13 ? (double)A[((int64_t)(d - 13) * 9 + 3 * ci + cj) * g.n_nodes + N]
                 : (double)A[((int64_t)(13 - d) * 9 + 3 * cj + ci) * g.n_nodes + M];
}

template <int R>
__global__ __launch_bounds__(1024) void k_mg_dense_inverse(const float* __restrict__ A, MgGrid g,
                                                           int n, float* __restrict__ inv) {
  // pivot row / column, double-buffered by pivot parity: one barrier per pivot suffices, since
  // the buffer written for pivot k + 2 was last read before every thread passed pivot k + 1's
  __shared__ double rowb[2][32 * R], colb[2][32 * R], dg[32 * R];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  double m[R][R];
#pragma unroll
  for (int a = 0; a < R; ++a)
#pragma unroll
    for (int b = 0; b < R; ++b) {
      const int i = ty + 32 * a, j = tx + 32 * b;
      m[a][b] = (i < n && j < n) ? mg_dense_entry(A, g, i, j) : (i == j ? 1.0 : 0.0);
      if (i == j) dg[i] = m[a][b];
    }
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    const int ka = k >> 5, kr = k & 31;
    double* rowv = rowb[k & 1];
    double* colv = colb[k & 1];
    // pivot row and column (read before any entry of this step changes); the register array
    // is indexed with compile-time indices only (ka is uniform, so the tests are scalar)
#pragma unroll
    for (int a = 0; a < R; ++a)
      if (a == ka && ty == kr) {
#pragma unroll
        for (int b = 0; b < R; ++b) rowv[tx + 32 * b] = m[a][b];
      }
#pragma unroll
    for (int b = 0; b < R; ++b)
      if (b == ka && tx == kr) {
#pragma unroll
        for (int a = 0; a < R; ++a) colv[ty + 32 * a] = m[a][b];
      }
    __syncthreads();
    const double p = colv[k], dk = dg[k];
    const bool dec = !(p > 1e-13 * dk);
    const double ip = dec ? 0.0 : 1.0 / p;
    const double dval = dec ? (dk > 0.0 ? 1.0 / dk : 1.0) : ip;
    // every entry: a_ij -= (a_ik / p) a_kj (this zeroes row and column k), then the owners of
    // row k, column k and the pivot overwrite them
#pragma unroll
    for (int a = 0; a < R; ++a) {
      const double c = colv[ty + 32 * a] * ip;
#pragma unroll
      for (int b = 0; b < R; ++b) m[a][b] = fma(-c, rowv[tx + 32 * b], m[a][b]);
    }
#pragma unroll
    for (int a = 0; a < R; ++a)
      if (a == ka && ty == kr) {
#pragma unroll
        for (int b = 0; b < R; ++b) m[a][b] = rowv[tx + 32 * b] * ip;
      }
#pragma unroll
    for (int b = 0; b < R; ++b)
      if (b == ka && tx == kr) {
#pragma unroll
        for (int a = 0; a < R; ++a) m[a][b] = -colv[ty + 32 * a] * ip;
      }
#pragma unroll
    for (int a = 0; a < R; ++a)
#pragma unroll
      for (int b = 0; b < R; ++b)
        if (a == ka && b == ka && ty == kr && tx == kr) m[a][b] = dval;
  }
#pragma unroll
  for (int a = 0; a < R; ++a)
#pragma unroll
    for (int b = 0; b < R; ++b) {
      const int i = ty + 32 * a, j = tx + 32 * b;
      if (i < n && j < n) inv[(int64_t)i * n + j] = (float)m[a][b];
    }
}

// Level 0 (matrix-free, FP32 s and Ke0, LDS tile): x0 = wd*r, t = r - K x0; r is the FP64
// CG residual.
__global__ __launch_bounds__(kBlock) void k_mg0_pre(const double* __restrict__ r, GridDims g,
                                                    TileGrid tg, const float* __restrict__ s,
                                                    const float* __restrict__ Ke,
                                                    const uint8_t* __restrict__ fixcnt,
                                                    const double* __restrict__ P,
                                                    const float* __restrict__ wd,
                                                    float* __restrict__ x0,
                                                    float* __restrict__ t, const int* done) {
  if (*done) return;
  __shared__ TileSmem<float> sm;
  int i0, j0, k0;
  if (!tile_origin(tg, i0, j0, k0)) return;
  int li, lj, lk;
  tile_thread(li, lj, lk);
  const int i = i0 + li, j = j0 + lj, k = k0 + lk;
  const bool own = i <= g.nx && j <= g.ny && k <= g.nz;
  const int64_t n = ((int64_t)i * g.nny + j) * g.nnz + k;
  // the epilogue's own-node loads are issued with the tile's, so their latency hides behind
  // the tile load and the operator instead of following them
  float rv[3] = {0.f, 0.f, 0.f}, fc[3] = {0.f, 0.f, 0.f};
  if (own) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      rv[c] = (float)r[3 * n + c];
      fc[c] = (float)fixcnt[3 * n + c];
    }
  }
  tile_load(sm, g, i0, j0, k0, [&](int64_t q) { return wd[q] * (float)r[q]; }, s);
  __syncthreads();
  if (!own) return;
  float y[3], xo[3];
  tile_apply(sm, li, lj, lk, Ke, y, xo);
  const float pen = (float)(*P);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int64_t q = 3 * n + c;
    y[c] = fmaf(fc[c] * pen, xo[c], y[c]);
    x0[q] = xo[c];
    t[q] = fc[c] > 0.f ? 0.f : rv[c] - y[c];  // clamped residuals are not restricted
  }
}

// Level 0 post-smoothing into the FP64 preconditioned residual z = x + wd*(r - K x), with
// block partials of z.r and z.q (the flexible-CG coefficients; q may be null).
__global__ __launch_bounds__(kBlock) void k_mg0_post(const double* __restrict__ r, GridDims g,
                                                     TileGrid tg, const float* __restrict__ s,
                                                     const float* __restrict__ Ke,
                                                     const uint8_t* __restrict__ fixcnt,
                                                     const double* __restrict__ P,
                                                     const float* __restrict__ wd,
                                                     const float* __restrict__ x,
                                                     double* __restrict__ z,
                                                     const double* __restrict__ q,
                                                     double* partials, int stride,
                                                     const int* done) {
  if (*done) return;
  __shared__ TileSmem<float> sm;
  int i0, j0, k0;
  const bool has = tile_origin(tg, i0, j0, k0);
  int li, lj, lk;
  tile_thread(li, lj, lk);
  const int i = i0 + li, j = j0 + lj, k = k0 + lk;
  const bool own = has && i <= g.nx && j <= g.ny && k <= g.nz;
  const int64_t n = ((int64_t)i * g.nny + j) * g.nnz + k;
  // own-node epilogue loads issued with the tile's (their latency hides behind the operator)
  double rv[3] = {0.0, 0.0, 0.0}, qv[3] = {0.0, 0.0, 0.0};
  float wv[3] = {0.f, 0.f, 0.f}, fc[3] = {0.f, 0.f, 0.f};
  if (own) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      rv[c] = r[3 * n + c];
      wv[c] = wd[3 * n + c];
      fc[c] = (float)fixcnt[3 * n + c];
      if (q) qv[c] = q[3 * n + c];
    }
  }
  if (has) tile_load(sm, g, i0, j0, k0, [&](int64_t qq) { return x[qq]; }, s);
  __syncthreads();
  double acc[2] = {0.0, 0.0};
  if (own) {
    float y[3], xo[3];
    tile_apply(sm, li, lj, lk, Ke, y, xo);
    const float pen = (float)(*P);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t qq = 3 * n + c;
      y[c] = fmaf(fc[c] * pen, xo[c], y[c]);
      // wd == 0 marks a DOF no element touches (zero row): its correction stays 0
      const double zv = wv[c] == 0.f
                            ? 0.0
                            : (double)xo[c] + (double)wv[c] * (rv[c] - (double)y[c]);
      z[qq] = zv;
      acc[0] = fma(zv, rv[c], acc[0]);
      acc[1] = fma(zv, qv[c], acc[1]);
    }
  }
  store_partials<2>(acc, partials, stride);
}

// ---------------------------------------------------------------------------------------
// Flexible CG around the V-cycle (FP64 vectors, operator and reductions)
// ---------------------------------------------------------------------------------------
struct FcgState {
  double rz, rr, rhs2, thresh, beta, pq;
  int it, done, converged, maxit;
};

// rr0 = |r0|^2 and bb = |b|^2 from device scalars.
__global__ void k_fcg_start(FcgState* st, const double* rr, const double* bb, double tol,
                            int maxit) {
  if (threadIdx.x || blockIdx.x) return;
  FcgState s;
  s.rhs2 = *bb;
  s.thresh = fmax(tol * tol * s.rhs2, 2.2250738585072014e-308);
  s.rr = *rr;
  s.rz = 0.0;
  s.beta = 0.0;
  s.pq = 0.0;
  s.it = 0;
  s.maxit = maxit;
  const int conv = (s.rhs2 == 0.0) || (s.rr < s.thresh);
  s.converged = conv;
  s.done = conv || maxit <= 0;
  *st = s;
}

// p = z + beta p_old (formed once per node in the LDS tile); q = K p (FP64); partials p.q.
// Held to 64 VGPRs (8 waves per SIMD, no spill): with its 19 KB of LDS, 8 blocks fit per CU, so
// the 120x60x60 grid (1,984 tiles) is resident in one round instead of 7 blocks per CU and a
// tail.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_fcg_apply(
    const double* __restrict__ z, const double* __restrict__ pold, double* __restrict__ pnew,
    double* __restrict__ q, const double* __restrict__ s, const double* __restrict__ Ke,
    const uint8_t* __restrict__ fixcnt, const double* __restrict__ P, GridDims g, TileGrid tg,
    const FcgState* st, double* partials) {
  if (st->done) return;
  const double beta = st->beta;
  __shared__ TileSmem<double> sm;
  int i0, j0, k0;
  const bool has = tile_origin(tg, i0, j0, k0);
  int li, lj, lk;
  tile_thread(li, lj, lk);
  const int i = i0 + li, j = j0 + lj, k = k0 + lk;
  const bool own = has && i <= g.nx && j <= g.ny && k <= g.nz;
  const int64_t n = ((int64_t)i * g.nny + j) * g.nnz + k;
  double fc[3] = {0.0, 0.0, 0.0};  // issued with the tile's loads
  if (own) {
#pragma unroll
    for (int c = 0; c < 3; ++c) fc[c] = (double)fixcnt[3 * n + c];
  }
  if (has)
    tile_load(sm, g, i0, j0, k0, [&](int64_t qq) { return fma(beta, pold[qq], z[qq]); }, s);
  __syncthreads();
  double acc[1] = {0.0};
  if (own) {
    double yv[3], pv[3];
    tile_apply(sm, li, lj, lk, Ke, yv, pv);
    const double pen = *P;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      yv[c] = fma(fc[c] * pen, pv[c], yv[c]);
      pnew[3 * n + c] = pv[c];
      q[3 * n + c] = yv[c];
      acc[0] = fma(pv[c], yv[c], acc[0]);
    }
  }
  store_partials<1>(acc, partials, gridDim.x);
}

// alpha = rz / p.q; x += alpha p; r -= alpha q; partials |r|^2.
__global__ __launch_bounds__(kBlock) void k_fcg_update(double* __restrict__ x,
                                                       double* __restrict__ r,
                                                       const double* __restrict__ p,
                                                       const double* __restrict__ q, int64_t n,
                                                       const FcgState* st,
                                                       const double* __restrict__ apart,
                                                       int nba, double* __restrict__ upart) {
  double pq[1];
  reduce_partials<1>(apart, nba, nba, pq);  // before the flag test: the loads overlap it
  if (st->done) return;
  const double a = st->rz / pq[0];
  double acc[1] = {0.0};
  AMT_GRID_STRIDE(i, n) {
    x[i] = fma(a, p[i], x[i]);
    const double rv = fma(-a, q[i], r[i]);
    r[i] = rv;
    acc[0] = fma(rv, rv, acc[0]);
  }
  store_partials<1>(acc, upart, gridDim.x);
}

// Level-0 pre-smoothing with the flexible-CG update folded in (one launch for k_fcg_update +
// k_mg0_pre): alpha = rz / p.q from the operator launch's partials; the tile loads r and q and
// forms r' = r - alpha q (k_fcg_update's fma) at every halo node, so that x0 = wd r' and
// t = r' - K x0 are k_mg0_pre's. The tile's own nodes write r' to rnew, a second buffer, because
// other blocks still read r in their halos. They also write x += alpha p and block partials of
// |r'|^2, one per block. The halo loads follow the partials behind alpha (their values are
// transformed as they arrive) and the own-node loads follow the element loop, so few
// registers stay live and 8 waves fit per SIMD (no spill at the 64-VGPR limit).
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_mg0_pre_upd(
    const double* __restrict__ r, double* __restrict__ rnew, const double* __restrict__ qv,
    double* __restrict__ x, const double* __restrict__ pv, GridDims g, TileGrid tg,
    const float* __restrict__ s, const float* __restrict__ Ke,
    const uint8_t* __restrict__ fixcnt, const double* __restrict__ P,
    const float* __restrict__ wd, float* __restrict__ x0, float* __restrict__ t,
    const FcgState* st, const double* __restrict__ apart, int nba, double* __restrict__ upart) {
  __shared__ TileSmem<float> sm;
  int i0 = 0, j0 = 0, k0 = 0;
  const bool has = tile_origin(tg, i0, j0, k0);
  int li, lj, lk;
  tile_thread(li, lj, lk);
  const int i = i0 + li, j = j0 + lj, k = k0 + lk;
  const bool own = has && i <= g.nx && j <= g.ny && k <= g.nz;
  const int64_t n = ((int64_t)i * g.nny + j) * g.nnz + k;
  float fc[3] = {0.f, 0.f, 0.f};
  if (own) {
#pragma unroll
    for (int c = 0; c < 3; ++c) fc[c] = (float)fixcnt[3 * n + c];
  }
  double pq[1];
  reduce_partials<1>(apart, nba, nba, pq);
  if (st->done) return;  // uniform over the grid, before any barrier
  const double a = st->rz / pq[0];
  if (has)
    tile_load(sm, g, i0, j0, k0,
              [&](int64_t q) { return wd[q] * (float)fma(-a, qv[q], r[q]); }, s);
  __syncthreads();
  double acc[1] = {0.0};
  if (own) {
    float y[3], xs[3];
    tile_apply(sm, li, lj, lk, Ke, y, xs);
    double ro[3], qo[3], xo[3], po[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      ro[c] = r[3 * n + c];
      qo[c] = qv[3 * n + c];
      xo[c] = x[3 * n + c];
      po[c] = pv[3 * n + c];
    }
    const float pen = (float)(*P);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t q = 3 * n + c;
      const double rn = fma(-a, qo[c], ro[c]);
      rnew[q] = rn;
      x[q] = fma(a, po[c], xo[c]);
      acc[0] = fma(rn, rn, acc[0]);
      y[c] = fmaf(fc[c] * pen, xs[c], y[c]);
      x0[q] = xs[c];
      t[q] = fc[c] > 0.f ? 0.f : (float)rn - y[c];  // clamped residuals are not restricted
    }
  }
  store_partials<1>(acc, upart, gridDim.x);
}

// One block after the V-cycle: convergence test on |r|^2, then the flexible (Polak-Ribiere /
// Notay) beta = -(z.q)/(p.q) and rz = z.r. first=1: only rz from the first V-cycle.
__global__ __launch_bounds__(kBlock) void k_fcg_scalar(FcgState* st,
                                                       const double* __restrict__ upart,
                                                       int nbu, const double* __restrict__ apart,
                                                       int nba, const double* __restrict__ zpart,
                                                       int nbz, int zstride, int first,
                                                       int* done_flag) {
  // z.r, z.q, |r|^2 and p.q in one pass: every thread folds its strided partials of the four
  // sums in ascending order (the order reduce_partials uses), four strides' loads in flight,
  // then one block reduction; all before the flag test, so the loads overlap it
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  const int nmax = max(nbz, max(nbu, nba));
  for (int b0 = threadIdx.x; b0 < nmax; b0 += 4 * kBlock) {
    double x[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int b = b0 + u * kBlock;
      x[u][0] = b < nbz ? zpart[b] : 0.0;
      x[u][1] = b < nbz ? zpart[(size_t)zstride + b] : 0.0;
      x[u][2] = (!first && b < nbu) ? upart[b] : 0.0;
      x[u][3] = (!first && b < nba) ? apart[b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int c = 0; c < 4; ++c) v[c] += x[u][c];
  }
  double tot[4];
  block_reduce<4>(v, tot);
  const double zr[2] = {tot[0], tot[1]};
  FcgState s = *st;
  if (s.done) {
    if (threadIdx.x == 0) *done_flag = 1;
    return;
  }
  if (first) {
    if (threadIdx.x == 0) {
      s.rz = zr[0];
      s.beta = 0.0;
      *st = s;
    }
    return;
  }
  const double rr[1] = {tot[2]}, pq[1] = {tot[3]};
  if (threadIdx.x != 0) return;
  s.rr = rr[0];
  s.pq = pq[0];
  if (rr[0] < s.thresh) {
    s.done = 1;
    s.converged = 1;
  } else {
    s.beta = -zr[1] / pq[0];
    s.rz = zr[0];
    s.it += 1;
    if (s.it >= s.maxit) s.done = 1;
  }
  *st = s;
  *done_flag = s.done;
}

}  // namespace amt
