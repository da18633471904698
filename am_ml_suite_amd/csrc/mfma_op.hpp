// mfma_op.hpp — the fine-grid operator y = sum_e s_e Ke0 x_e on the matrix cores (gfx950
// v_mfma_f32_16x16x16_bf16), for the FP32 parts of the solvers.
//
// Element-batched formulation. A 256-thread block owns the same 4x4x16 node tile as the vector
// operator (tile_op.hpp) and the 5x5x17 = 425 elements touching it. The element products
// Y = Ke0 X (X: 24 x 425, the elements' nodal vectors) are one GEMM of 24 x 24 by 24 x 425:
// with 16 x 16 x 16 tiles that is 2 (row halves) x 2 (k) x 27 (element groups) MFMAs, about
// 1/16 of the 576 FMAs per node of the vector operator in issue slots. The halo is staged in
// LDS, the products are scaled by s_e and written to LDS, and each thread then sums the rows of
// its node from its (up to) eight elements in a fixed order.
//
// Accuracy: x and Ke0 are each split into two BF16 terms, v = hi + lo with hi = bf16(v) and
// lo = bf16(v - hi), and Y = Kh Xh + Kh Xl + Kl Xh with FP32 accumulation (the Kl Xl term,
// ~2^-16 relative, is dropped): about 1e-5 relative error against the FP64 operator, i.e.
// close to the FP32 vector operator, at three MFMAs per product tile.
//
// MFMA operand layout (wave64, 16x16x16 bf16; checked against the CPU operator by the tests):
//   A (M x K): lane l holds A[l % 16][4 (l / 16) + e]    -> Ke0[m][k]
//   B (K x N): lane l holds B[4 (l / 16) + e][l % 16]    -> X[k][element]
//   D (M x N): lane l holds D[4 (l / 16) + e][l % 16]    -> Y[m][element]
#pragma once

#include "device.hpp"
#include "kernels.hpp"
#include "multigrid.hpp"

namespace amt {

typedef short mf_bf16x4 __attribute__((ext_vector_type(4)));
typedef float mf_f32x4 __attribute__((ext_vector_type(4)));

// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t mf_bf16(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float mf_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// The element products are formed in two halves of 12 rows (corners 0-3, then 4-7), each from
// one 16-row MFMA tile starting at row 12h: the same 2 x 2 x 3 MFMAs per element group as two
// fixed row tiles, but the LDS product buffer holds 12 rows per element instead of 24 (22 KB
// instead of 43 KB per block: 5 blocks per CU instead of 3).
constexpr int kYS = 13;  // LDS row stride (floats) of one element's 12 products (odd: banks)
constexpr int kEGroups = (kETile + 15) / 16;  // 27 groups of 16 elements

struct MfmaSmem {
  uint16_t xh[kHalo * 3], xl[kHalo * 3];
  float s[kETile];
  float y[kETile * kYS];
};

// The lane's Ke0 fragments: [half][k step], high and low BF16 parts; half h holds rows
// 12h .. 12h + 15 (rows past 23 are zero).
struct KeFrag {
  mf_bf16x4 h[2][2], l[2][2];
};

// Kh / Kl: 24 x 24 BF16 high and low parts of Ke0 (row-major).
__device__ __forceinline__ void load_kefrag(KeFrag& f, const uint16_t* __restrict__ Kh,
                                            const uint16_t* __restrict__ Kl) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt)
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int m = mt * 12 + (lane & 15), k = ks * 16 + 4 * (lane >> 4) + e;
        const bool ok = m < 24 && k < 24;
        f.h[mt][ks][e] = ok ? (short)Kh[m * 24 + k] : (short)0;
        f.l[mt][ks][e] = ok ? (short)Kl[m * 24 + k] : (short)0;
      }
}

// Halo node values f(dof) (0 outside the grid) split into BF16 hi/lo, and element scales
// (all global loads issued first, as tile_load).
template <typename F>
__device__ __forceinline__ void mfma_tile_load(MfmaSmem& sm, const GridDims& g, int i0, int j0,
                                               int k0, F f, const float* __restrict__ s) {
  float v[kNPer][3], sv[kSPer];
  tile_gather(v, sv, g, i0, j0, k0, f, s);
#pragma unroll
  for (int t = 0; t < kNPer; ++t) {
    const int node = threadIdx.x + t * kBlock;
    if (node < kHalo) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const uint16_t h = mf_bf16(v[t][c]);
        sm.xh[3 * node + c] = h;
        sm.xl[3 * node + c] = mf_bf16(v[t][c] - mf_f32(h));
      }
    }
  }
#pragma unroll
  for (int t = 0; t < kSPer; ++t) {
    const int q = threadIdx.x + t * kBlock;
    if (q < kETile) sm.s[q] = sv[t];
  }
}

// Rows 12h .. 12h + 11 of the element products s_e * Ke0 x_e of the tile's 425 elements into
// sm.y (after mfma_tile_load and a barrier; the caller syncs again before reading sm.y).
__device__ __forceinline__ void mfma_tile_products(MfmaSmem& sm, const KeFrag& kf, int half) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int col = lane & 15, kq = lane >> 4;
  // LDS offsets (relative to the element's first node) of the lane's 8 k values
  int koff[2][4];
  bool kok[2][4];
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int k = ks * 16 + 4 * kq + e;
      const int a = k / 3, c = k - 3 * (k / 3);
      const int ox = (a ^ (a >> 1)) & 1, oy = (a >> 1) & 1, oz = (a >> 2) & 1;  // kOX/kOY/kOZ
      kok[ks][e] = k < 24;
      koff[ks][e] = 3 * ((ox * kHJ + oy) * kHK + oz) + c;
    }
  for (int grp = wave; grp < kEGroups; grp += kBlock / 64) {
    const int en = grp * 16 + col;
    const bool ev = en < kETile;
    const int ei = en / (kEJ * kEK), ej = (en / kEK) % kEJ, ek = en % kEK;
    const int base = 3 * ((ei * kHJ + ej) * kHK + ek);
    mf_bf16x4 bh[2], bl[2];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const bool ok = ev && kok[ks][e];
        bh[ks][e] = ok ? (short)sm.xh[base + koff[ks][e]] : (short)0;
        bl[ks][e] = ok ? (short)sm.xl[base + koff[ks][e]] : (short)0;
      }
    const float se = ev ? sm.s[en] : 0.f;
    mf_f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const mf_bf16x4 kh = half ? kf.h[1][ks] : kf.h[0][ks];
      const mf_bf16x4 kl = half ? kf.l[1][ks] : kf.l[0][ks];
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(kh, bh[ks], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(kh, bl[ks], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(kl, bh[ks], acc, 0, 0, 0);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int m = 4 * kq + e;  // row 12 half + m of Ke0
      if (ev && m < 12) sm.y[en * kYS + m] = se * acc[e];
    }
  }
}

// The thread's node (tile position li, lj, lk): y[c] += its rows from the elements where it is
// corner 4 half .. 4 half + 3 (over both halves: all 8 elements, in corner order).
__device__ __forceinline__ void mfma_node_sum(const MfmaSmem& sm, int li, int lj, int lk,
                                              int half, float (&y)[3]) {
  const int eb = ((li + 1) * kEJ + (lj + 1)) * kEK + (lk + 1);
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int a = 4 * half + u;
    const float* row = sm.y + (eb - (kOX[a] * kEJ + kOY[a]) * kEK - kOZ[a]) * kYS + 3 * u;
    y[0] += row[0];
    y[1] += row[1];
    y[2] += row[2];
  }
}

// y = sum_e s_e Ke0 x_e at the thread's node, after mfma_tile_load and a barrier: both halves
// of the products, each summed into y (every thread of the block calls it; has: the block
// has a tile).
__device__ __forceinline__ void mfma_tile_operator(MfmaSmem& sm, const KeFrag& kf, bool has,
                                                   int li, int lj, int lk, float (&y)[3]) {
  y[0] = y[1] = y[2] = 0.f;
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    if (half) __syncthreads();  // every node sum of the first half is done
    if (has) mfma_tile_products(sm, kf, half);
    __syncthreads();
    if (has) mfma_node_sum(sm, li, lj, lk, half, y);
  }
}

// y = K x (+ penalty rows when P != null), FP32 in and out.
__global__ __launch_bounds__(kBlock) void k_apply_mfma(const float* __restrict__ x,
                                                       float* __restrict__ y,
                                                       const float* __restrict__ s,
                                                       const uint16_t* __restrict__ Kh,
                                                       const uint16_t* __restrict__ Kl,
                                                       const uint8_t* __restrict__ fixcnt,
                                                       const double* __restrict__ P,
                                                       GridDims g, TileGrid tg) {
  __shared__ MfmaSmem sm;
  int i0, j0, k0;
  if (!tile_origin(tg, i0, j0, k0)) return;  // whole block: no barrier follows for it
  KeFrag kf;
  load_kefrag(kf, Kh, Kl);
  mfma_tile_load(sm, g, i0, j0, k0, [&](int64_t q) { return x[q]; }, s);
  __syncthreads();
  int li, lj, lk;
  tile_thread(li, lj, lk);
  float yv[3];
  mfma_tile_operator(sm, kf, true, li, lj, lk, yv);
  const int i = i0 + li, j = j0 + lj, k = k0 + lk;
  if (i > g.nx || j > g.ny || k > g.nz) return;
  const int64_t n = ((int64_t)i * g.nny + j) * g.nnz + k;
  const float pen = P ? (float)(*P) : 0.f;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int64_t q = 3 * n + c;
    y[q] = P ? fmaf((float)fixcnt[q] * pen, x[q], yv[c]) : yv[c];
  }
}

// PCG iteration t, first half, on the MFMA operator (as k_pcg_apply<float, float>).
__global__ __launch_bounds__(kBlock) void k_pcg_apply_mfma(
    int t, const float* __restrict__ z, const float* __restrict__ pold, float* __restrict__ pnew,
    float* __restrict__ q, const float* __restrict__ s, const uint16_t* __restrict__ Kh,
    const uint16_t* __restrict__ Kl, const uint8_t* __restrict__ fixcnt,
    const double* __restrict__ P, GridDims g, TileGrid tg, PcgState* S,
    const double* __restrict__ upart, int nbu, double* __restrict__ apart, int stride) {
  const PcgState cur = S[t & 1];
  if (cur.done) {
    if (blockIdx.x == 0 && threadIdx.x == 0) S[(t + 1) & 1] = cur;
    return;
  }
  double beta_d;
  const PcgState nx = pcg_step_state(cur, t, upart, nbu, stride, beta_d);
  if (blockIdx.x == 0 && threadIdx.x == 0) S[(t + 1) & 1] = nx;
  if (nx.done) return;
  const float beta = (float)beta_d;
  __shared__ MfmaSmem sm;
  int i0, j0, k0;
  const bool has = tile_origin(tg, i0, j0, k0);
  KeFrag kf;
  load_kefrag(kf, Kh, Kl);
  if (has)
    mfma_tile_load(sm, g, i0, j0, k0, [&](int64_t qq) { return fmaf(beta, pold[qq], z[qq]); }, s);
  __syncthreads();
  int li, lj, lk;
  tile_thread(li, lj, lk);
  float yv[3];
  mfma_tile_operator(sm, kf, has, li, lj, lk, yv);
  const int i = i0 + li, j = j0 + lj, k = k0 + lk;
  double acc[1] = {0.0};
  if (has && i <= g.nx && j <= g.ny && k <= g.nz) {
    const int64_t n = ((int64_t)i * g.nny + j) * g.nnz + k;
    const float pen = (float)(*P);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t qq = 3 * n + c;
      const float pv = fmaf(beta, pold[qq], z[qq]);
      const float yc = fmaf((float)fixcnt[qq] * pen, pv, yv[c]);
      pnew[qq] = pv;
      q[qq] = yc;
      acc[0] = fma((double)pv, (double)yc, acc[0]);
    }
  }
  store_partials<1>(acc, apart, stride);
}

// V-cycle level-0 pre-smoothing on the MFMA operator (as k_mg0_pre).
__global__ __launch_bounds__(kBlock) void k_mg0_pre_mfma(
    const double* __restrict__ r, GridDims g, TileGrid tg, const float* __restrict__ s,
    const uint16_t* __restrict__ Kh, const uint16_t* __restrict__ Kl,
    const uint8_t* __restrict__ fixcnt, const double* __restrict__ P,
    const float* __restrict__ wd, float* __restrict__ x0, float* __restrict__ t,
    const int* done) {
  if (*done) return;
  __shared__ MfmaSmem sm;
  int i0, j0, k0;
  if (!tile_origin(tg, i0, j0, k0)) return;  // whole block: no barrier follows for it
  KeFrag kf;
  load_kefrag(kf, Kh, Kl);
  mfma_tile_load(sm, g, i0, j0, k0, [&](int64_t q) { return wd[q] * (float)r[q]; }, s);
  __syncthreads();
  int li, lj, lk;
  tile_thread(li, lj, lk);
  float y[3];
  mfma_tile_operator(sm, kf, true, li, lj, lk, y);
  const int i = i0 + li, j = j0 + lj, k = k0 + lk;
  if (i > g.nx || j > g.ny || k > g.nz) return;
  const int64_t n = ((int64_t)i * g.nny + j) * g.nnz + k;
  const float pen = (float)(*P);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const int64_t q = 3 * n + c;
    const float xv = wd[q] * (float)r[q];
    y[c] = fmaf((float)fixcnt[q] * pen, xv, y[c]);
    x0[q] = xv;
    t[q] = fixcnt[q] ? 0.f : (float)r[q] - y[c];  // clamped residuals are not restricted
  }
}

// V-cycle level-0 post-smoothing on the MFMA operator (as k_mg0_post).
__global__ __launch_bounds__(kBlock) void k_mg0_post_mfma(
    const double* __restrict__ r, GridDims g, TileGrid tg, const float* __restrict__ s,
    const uint16_t* __restrict__ Kh, const uint16_t* __restrict__ Kl,
    const uint8_t* __restrict__ fixcnt, const double* __restrict__ P,
    const float* __restrict__ wd, const float* __restrict__ x, double* __restrict__ z,
    const double* __restrict__ q, double* partials, int stride, const int* done) {
  if (*done) return;
  __shared__ MfmaSmem sm;
  int i0, j0, k0;
  const bool has = tile_origin(tg, i0, j0, k0);
  KeFrag kf;
  load_kefrag(kf, Kh, Kl);
  if (has) mfma_tile_load(sm, g, i0, j0, k0, [&](int64_t qq) { return x[qq]; }, s);
  __syncthreads();
  int li, lj, lk;
  tile_thread(li, lj, lk);
  float y[3];
  mfma_tile_operator(sm, kf, has, li, lj, lk, y);
  const int i = i0 + li, j = j0 + lj, k = k0 + lk;
  double acc[2] = {0.0, 0.0};
  if (has && i <= g.nx && j <= g.ny && k <= g.nz) {
    const int64_t n = ((int64_t)i * g.nny + j) * g.nnz + k;
    const float pen = (float)(*P);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int64_t qq = 3 * n + c;
      const float xv = x[qq];
      y[c] = fmaf((float)fixcnt[qq] * pen, xv, y[c]);
      const double rv = r[qq];
      // wd == 0 marks a DOF no element touches (zero row): its correction stays 0
      const double zv = wd[qq] == 0.f ? 0.0 : (double)xv + (double)wd[qq] * (rv - (double)y[c]);
      z[qq] = zv;
      acc[0] = fma(zv, rv, acc[0]);
      if (q) acc[1] = fma(zv, q[qq], acc[1]);
    }
  }
  store_partials<2>(acc, partials, stride);
}

}  // namespace amt
