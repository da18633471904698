// device.hpp — data structures and __device__ building blocks shared by every translation
// unit with device code: grid descriptors, solver states, deterministic fence-free
// reductions, the hex8 node operator and its LDS tile, grid-stride loops, the filter
// stencil entry and the OC candidate arithmetic. Kernels live in kernels.hpp (single
// problem) and batch.hpp (many problems).
//
// Data layout in HBM (one problem):
//   nodal vectors  : double/float [3*N_n], DOF = 3*node + c, node z-fastest
//                    (element_kernel.cpp:186-216)
//   element scale s: [N_e] in C order (i*ny*nz + j*nz + k), so the eight elements around a
//                    node are three contiguous pairs along k
//   slot arrays    : rho, dc, dcf, design mask [N_e] in the optimizer's x-fastest slot order
//                    (topology_optimizer.hpp:218-232)
//
// Reductions are deterministic: each block reduces in a fixed tree order and writes one
// partial; the partials are summed in a fixed order by the next kernel in the stream.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace amt {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

// element_kernel.cpp:196-199: local node offsets of the hex8 element
__device__ constexpr int kOX[8] = {0, 1, 1, 0, 0, 1, 1, 0};
__device__ constexpr int kOY[8] = {0, 0, 1, 1, 0, 0, 1, 1};
__device__ constexpr int kOZ[8] = {0, 0, 0, 0, 1, 1, 1, 1};

struct GridDims {
  int nx, ny, nz;  // elements
  int nny, nnz;    // ny+1, nz+1
  int64_t n_nodes;
  int64_t n_elem;
};

// Device-resident PCG state. Two copies alternate by iteration parity: iteration t reads
// S[t&1] and block 0 writes S[(t+1)&1], so no kernel reads a slot it is writing.
struct PcgState {
  double rz, rr, rhs2, thresh;
  int it, done, converged, maxit;
};

struct OcState {
  double l1, l2, lmid, vf, n_design, change, vol_sum, pad;
  int passes, done, pad1, pad2;
  // passes whose outcome is known (k_oc_bounds): lambda + 1e-20 >= ff_hi -> volume vol_hi,
  // lambda + 1e-20 <= ff_lo -> volume vol_lo (ff_hi = inf, ff_lo = -inf: none known)
  double ff_hi, ff_lo, vol_hi, vol_lo;
};

// ---------------------------------------------------------------------------------------
// Deterministic, fence-free reductions
//
// A block reduces its values in a fixed tree order and stores one partial per value with a
// plain store. The partials are reduced by the NEXT kernel in the stream (every block of it,
// in the same fixed order, so all blocks obtain bitwise-identical scalars) or by a one-block
// finalize kernel. The kernel boundary makes the partials visible. No block issues an
// agent-scope fence: on gfx950 each such fence writes back the XCD's L2, which costs more
// than the whole reduction when the kernel has just written tens of MB.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Block-reduce NV values; every thread receives the block totals in out.
template <int NV, bool IS_MAX = false>
__device__ __forceinline__ void block_reduce(const double (&v)[NV], double (&out)[NV]) {
  __shared__ double red[NV][kWaves];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const double t = IS_MAX ? wave_max(v[q]) : wave_sum(v[q]);
    if (lane == 0) red[q][w] = t;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double t = red[q][0];
#pragma unroll
    for (int k = 1; k < kWaves; ++k) t = IS_MAX ? fmax(t, red[q][k]) : t + red[q][k];
    out[q] = t;
  }
  __syncthreads();
}

// partials[q*stride + blockIdx.x] = block total of v[q]. All threads of the block call it.
template <int NV, bool IS_MAX = false>
__device__ __forceinline__ void store_partials(const double (&v)[NV], double* partials,
                                               int stride) {
  double t[NV];
  block_reduce<NV, IS_MAX>(v, t);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) partials[(size_t)q * stride + blockIdx.x] = t[q];
  }
}

// Block totals of many values (NV <= 64) at once. A wave reduce-scatter by recursive
// halving: at distance d each lane keeps d of its 2d running sums and adds the partner's copy
// of the same d, so after 32 + 16 + ... + 1 = 63 shuffles lane l holds the wave total of
// value l (instead of 6 shuffles per value); the waves' totals are then summed in wave order
// into tot[0..NV). All threads of the block call it; tot is valid after it returns.
template <int NV>
__device__ __forceinline__ void block_reduce_many(const double (&v)[NV], double* tot) {
  static_assert(NV <= 64, "block_reduce_many holds one value per lane");
  __shared__ double red[kWaves][64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double a[64];
#pragma unroll
  for (int i = 0; i < 64; ++i) a[i] = i < NV ? v[i] : 0.0;
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const bool up = lane & d;
#pragma unroll
    for (int i = 0; i < d; ++i) {
      const double send = up ? a[i] : a[d + i];
      const double keep = up ? a[d + i] : a[i];
      a[i] = keep + __shfl_xor(send, d, 64);
    }
  }
  red[w][lane] = a[0];
  __syncthreads();
  if (threadIdx.x < NV) {
    double t = red[0][threadIdx.x];
#pragma unroll
    for (int k = 1; k < kWaves; ++k) t += red[k][threadIdx.x];
    tot[threadIdx.x] = t;
  }
  __syncthreads();
}

// partials[q*stride + blockIdx.x] = block total of v[q], for many values (NV <= 64).
template <int NV>
__device__ __forceinline__ void store_partials_many(const double (&v)[NV], double* partials,
                                                    int stride) {
  __shared__ double tot[NV];
  block_reduce_many<NV>(v, tot);
  if (threadIdx.x < NV) partials[(size_t)threadIdx.x * stride + blockIdx.x] = tot[threadIdx.x];
}

// out[q] = fixed-order reduction of partials[q*stride + 0..nb). All threads of the block
// call it and all receive the result. Thread t folds partials t, t + kBlock, ... in ascending
// order; the loads of U strides and all NV values are issued before the first add, so a
// reduction over a few thousand partials waits on memory once or twice instead of once per
// value.
template <int NV, bool IS_MAX = false>
__device__ __forceinline__ void reduce_partials(const double* partials, int nb, int stride,
                                                double (&out)[NV]) {
  constexpr int U = NV == 1 ? 8 : (NV <= 5 ? 4 : 1);
  double v[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) v[q] = 0.0;
  int b = threadIdx.x;
  for (; b + (U - 1) * kBlock < nb; b += U * kBlock) {
    double x[U][NV];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int q = 0; q < NV; ++q) x[u][q] = partials[(size_t)q * stride + b + u * kBlock];
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int q = 0; q < NV; ++q) v[q] = IS_MAX ? fmax(v[q], x[u][q]) : v[q] + x[u][q];
  }
  for (; b < nb; b += kBlock) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const double x = partials[(size_t)q * stride + b];
      v[q] = IS_MAX ? fmax(v[q], x) : v[q] + x;
    }
  }
  if constexpr (NV > 8 && !IS_MAX) {
    __shared__ double tot[NV];
    block_reduce_many<NV>(v, tot);
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] = tot[q];
  } else {
    block_reduce<NV, IS_MAX>(v, out);
  }
}

template <int NV, bool IS_MAX = false>
__global__ __launch_bounds__(kBlock) void k_finalize(const double* __restrict__ partials,
                                                     int nb, int stride,
                                                     double* __restrict__ out) {
  double t[NV];
  reduce_partials<NV, IS_MAX>(partials, nb, stride, t);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) out[q] = t[q];
  }
}

// ---------------------------------------------------------------------------------------
// Matrix-free hex8 operator, node-centric gather.
//
// One thread owns one node n=(i,j,k) and its three rows. It holds the 27 neighbour nodal
// vectors in registers and, for each of the (up to) eight elements e around n (n at local
// position a in e), adds s_e * Ke0[3a+c, :] . x_e. Ke0 is read through uniform (scalar)
// loads. The sum over elements runs in a fixed order, so the result is deterministic.
// ---------------------------------------------------------------------------------------
template <typename T>
struct Nbr {
  T v[27][3];
};

__device__ __forceinline__ void node_decode(int64_t n, const GridDims& g, int& i, int& j,
                                            int& k) {
  k = (int)(n % g.nnz);
  const int64_t t = n / g.nnz;
  j = (int)(t % g.nny);
  i = (int)(t / g.nny);
}

template <typename T>
__device__ __forceinline__ void load_element_scales(const T* __restrict__ s, const GridDims& g,
                                                    int i, int j, int k, T (&s8)[8]) {
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int ei = i - kOX[a], ej = j - kOY[a], ek = k - kOZ[a];
    const bool ok = ei >= 0 && ei < g.nx && ej >= 0 && ej < g.ny && ek >= 0 && ek < g.nz;
    s8[a] = ok ? s[((int64_t)ei * g.ny + ej) * g.nz + ek] : T(0);
  }
}

// y_n = sum_a s8[a] * Ke0[3a.., :] x_e(a)
template <typename T, typename TK>
__device__ __forceinline__ void apply_node(const Nbr<T>& nb, const T (&s8)[8],
                                           const TK* __restrict__ Ke, T (&y)[3]) {
  y[0] = y[1] = y[2] = T(0);
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    T t0 = T(0), t1 = T(0), t2 = T(0);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int di = kOX[b] - kOX[a] + 1, dj = kOY[b] - kOY[a] + 1, dk = kOZ[b] - kOZ[a] + 1;
      const int m = di * 9 + dj * 3 + dk;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const T xv = nb.v[m][d];
        t0 = fma((T)Ke[(3 * a + 0) * 24 + 3 * b + d], xv, t0);
        t1 = fma((T)Ke[(3 * a + 1) * 24 + 3 * b + d], xv, t1);
        t2 = fma((T)Ke[(3 * a + 2) * 24 + 3 * b + d], xv, t2);
      }
    }
    y[0] = fma(s8[a], t0, y[0]);
    y[1] = fma(s8[a], t1, y[1]);
    y[2] = fma(s8[a], t2, y[2]);
  }
}

}  // namespace amt
#include "tile_op.hpp"
namespace amt {

template <typename T>
__global__ void k_fill(T* __restrict__ v, T val, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) v[i] = val;
}

// Grid-stride helper with a fixed grid: element i of thread t is t + q*stride.
#define AMT_GRID_STRIDE(i, n) \
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < (n); i += (int64_t)gridDim.x * kBlock)

struct StencilEntry {
  int dx, dy, dz, pad;
};

// std::max / std::min semantics: max(a,b) = (a<b)?b:a, min(a,b) = (b<a)?b:a (NaN rules kept)
__device__ __forceinline__ double std_max(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double std_min(double a, double b) { return (b < a) ? b : a; }

__device__ __forceinline__ double oc_candidate(double r, double dcf, double lmid, double move,
                                               double rho_min) {
  const double Be = sqrt(-dcf / (lmid * 1.0 + 1e-20));
  const double c = r * Be;
  return std_max(rho_min, std_max(r - move, std_min(1.0, std_min(r + move, c))));
}

// Midpoint of bisection-tree node nd (heap order: node 1 = [l1,l2], child 2j = left half where
// l2 := mid, 2j+1 = right half where l1 := mid), with the reference's (l1+l2)/2 arithmetic
// along the path, so it equals the lambda a sequential bisection would evaluate there.
__device__ __forceinline__ double oc_tree_mid(double l1, double l2, int nd) {
  int depth = 31 - __clz(nd);
  for (int b = depth - 1; b >= 0; --b) {
    const double m = 0.5 * (l1 + l2);
    if ((nd >> b) & 1)
      l1 = m;
    else
      l2 = m;
  }
  return 0.5 * (l1 + l2);
}

}  // namespace amt
