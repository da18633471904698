// unet.hpp — BF16 3D U-Net kernels (gfx950): MFMA implicit-GEMM 3x3x3 convolution forward and
// data gradient, MFMA weight gradient (deterministic ordered sums), pooling/upsampling, ReLU
// backward, Adam.
//
// Tensors are channels-last on a padded voxel grid (Px, Py, Pz all multiples of 8, z fastest):
// T[((i*Py + j)*Pz + k) * C + c], BF16 (raw 16-bit patterns), channel counts multiples of 16.
// Convolution weights are stored per tap as Wt[tap][co][ci] (taps in dz-fastest order
// (di+1)*9 + (dj+1)*3 + (dk+1)), so the MFMA B operand of 4 consecutive input channels is one
// 8-byte load. Accumulation is FP32.
//
// The convolutions run on v_mfma_f32_16x16x32_bf16 (wave64; fragment layouts below), checked
// against a CPU convolution by tests/test_gpu_unet.py.
#pragma once

#include "unet_api.hpp"

namespace amt {

typedef float f32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------------------
// Convolution layers (3x3x3 with zero padding):
//   y[v][co] = act( bias[co] + sum_tap sum_ci W[tap][co][ci] * x[v + off(tap)][ci] )
// Input channels come from x1 (c1 channels) then x2 (c2 channels, may be 0): the U-Net skip
// concatenation is read in place. Output channels [0, ysplit) go to y1 (row stride ysplit),
// [ysplit, cout) to y2 (row stride cout - ysplit) — the data gradient of a concatenated input.
// accum: add to the output instead of overwriting.
//
// Implicit GEMM on v_mfma_f32_16x16x32_bf16 (wave64): M = output channels (16 per tile), N =
// voxels (16 per tile), K = (tap, input channel) in chunks of 32, each lane group of 16 lanes
// taking 8 consecutive channels of one tap, so every A and B fragment is one 16-byte load. The
// pairing of K between the A and B fragments is the same lane slot on both sides, which is all
// the product needs.
//   A fragment: lane l holds A[l % 16][8 (l / 16) .. + 7]; B: B[8 (l / 16) .. + 7][l % 16];
//   D: lane l holds D[4 (l / 16) + e][l % 16], e = 0..3 (row = channel, column = voxel).
// ---------------------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
union Frag8 {
  uint4 u;
  bf16x8 v;
};

// ---------------------------------------------------------------------------------------
// Brick convolution (the operation and arguments above). A block computes a 4 x 4 x 16 voxel
// brick (16 M tiles of 16 voxels along z) for one 16-channel output tile. Wave w takes i-planes
// 2 (w / 2) + {0, 1} and j-rows 2 (w % 2) + {0, 1} of the brick, its M tile m being plane
// 2 (w / 2) + m / 2, row 2 (w % 2) + m % 2, so a 2x2x2 pool window lies in two neighbouring
// lanes. Input channels go in groups of 16: the group's halo brick (6 x 6 x 18 voxels) is
// staged in LDS as two 8-channel halves ([half][voxel][8], so the 8 lanes of an LDS access read
// 128 contiguous bytes), and the 27 taps run from LDS, two taps per 16x16x32 MFMA (K = 2 taps x
// 16 channels; lane group g takes tap 2c + g / 2, half g % 2). The weight fragments of the
// channel tile for the group (14 chunks) sit in registers. The weights are the MFMA's A operand
// and the voxels its B operand, so lane (g, r) ends with output channels 4 g .. 4 g + 3 of
// voxel r of each M tile: one 8-byte store per tile.
// Grid: x = blocks walking the bricks (block x takes bricks x, x + gridDim.x, ...: the next
// brick's halo is loaded while the current one's MFMAs run, and with a single input group per
// brick the weight fragments are loaded once), y = output-channel tiles, z = splits of the
// input-channel groups (every split has at least one group). With one split the block applies
// bias, accumulation and ReLU itself (and the fused outputs below); with several, each writes
// FP32 partial sums part[z][v][co] and k_conv_finish adds them in split order (a completion
// counter with device-scope fences in this kernel measured twice as slow: the fences write back
// and invalidate L2). up1: x1 is at half resolution (dimensions d / 2) and is read as its
// nearest-neighbour upsampling, so the decoder's upsampled input is never written.
// ---------------------------------------------------------------------------------------
constexpr int kBrickHalo = 6 * 6 * 18;
// the second channel half starts 64 bytes past a multiple of 128 B: the 8 lanes of a 128-byte
// LDS store phase (4 voxels x 2 halves) then hit disjoint banks
constexpr int kBrickHalf = kBrickHalo + 4;

typedef short s16x4 __attribute__((ext_vector_type(4)));

// Outputs fused into an unsplit brick convolution's epilogue (null pointers: none). Both read
// the layer's BF16 outputs as stored:
//   pool: their 2x2x2 average pool (dimensions d / 2, all even; cout channels): the window's
//         eight BF16 values summed in FP32 in (di, dj, dk) order, dk fastest, times 1/8, rounded
//         to BF16 (the same sum, bit for bit, as k_conv_finish_pool's);
//   wh, bh, yh: the pointwise head (16 -> 16 channels, W[co][ci] BF16, FP32 bias, no
//         activation; cout = 16) on a 16x16x16 MFMA, written to yh.
struct ConvFuse {
  uint16_t* pool;
  const uint16_t* wh;
  const float* bh;
  uint16_t* yh;
};

__device__ __forceinline__ float bf_lo(uint32_t u) { return __uint_as_float(u << 16); }
__device__ __forceinline__ float bf_hi(uint32_t u) { return __uint_as_float(u & 0xFFFF0000u); }
__device__ __forceinline__ uint32_t bf_pack(float a, float b) {
  return (uint32_t)f2bf(a) | ((uint32_t)f2bf(b) << 16);
}

__global__ __launch_bounds__(256) void k_conv_brick(
    const uint16_t* __restrict__ x1, int c1, const uint16_t* __restrict__ x2, int c2,
    const uint16_t* __restrict__ W, const float* __restrict__ bias, uint16_t* __restrict__ y1,
    uint16_t* __restrict__ y2, int ysplit, int relu, int accum, VoxDims d, int cout,
    float* __restrict__ part, int up1, ConvFuse fz) {
  __shared__ uint4 sx[2 * kBrickHalf];  // [half][voxel] 8 channels each
  const int py1 = up1 ? d.py / 2 : d.py, pz1 = up1 ? d.pz / 2 : d.pz;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, r = lane & 15;
  const int wi = 2 * (w >> 1), wj = 2 * (w & 1);  // the wave's first plane and row in the brick
  const int nbk = (d.pz + 15) / 16, nbj = (d.py + 3) / 4;
  const int nbricks = ((d.px + 3) / 4) * nbj * nbk;
  const int cin = c1 + c2, ngr = cin / 16;
  const int gpk = (ngr + gridDim.z - 1) / gridDim.z;
  const int gbeg = blockIdx.z * gpk, gend = min(ngr, gbeg + gpk);
  const int per = gend - gbeg;  // input-channel groups per brick for this split
  const int co0 = blockIdx.y * 16 + 4 * g;  // the lane's four output channels
  // the block's output channels are fixed: their biases are read once (a read per output value
  // would wait on the next brick's halo loads, which share the load counter)
  float bco[4], bh[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int e = 0; e < 4; ++e) bco[e] = bias ? bias[co0 + e] : 0.f;
  s16x4 wh = {0, 0, 0, 0};  // head A fragment: W[r][4 g .. 4 g + 3]
  if (fz.yh) {
    wh = *(const s16x4*)(fz.wh + 16 * r + 4 * g);
#pragma unroll
    for (int e = 0; e < 4; ++e) bh[e] = fz.bh[4 * g + e];
  }
  // work items (brick, group), brick-major; block x takes bricks x, x + gridDim.x, ...
  const int my_bricks = blockIdx.x < nbricks ? (nbricks - 1 - blockIdx.x) / gridDim.x + 1 : 0;
  const int items = per > 0 ? my_bricks * per : 0;
  auto brick_of = [&](int t, int& i0, int& j0, int& k0) {
    const int b = blockIdx.x + (t / per) * gridDim.x;
    const int bk = b % nbk, bj = (b / nbk) % nbj, bi = b / (nbk * nbj);
    i0 = 4 * bi;
    j0 = 4 * bj;
    k0 = 16 * bk;
  };
  // the halo brick of item t: every load issued (consumed an item later)
  auto load_halo = [&](int t, uint4 (&hv)[6]) {
    int i0, j0, k0;
    brick_of(t, i0, j0, k0);
    const int gr = gbeg + t % per;
#pragma unroll
    for (int u = 0; u < 6; ++u) {
      const int tt = threadIdx.x + 256 * u;
      const int vx = tt >> 1, half = tt & 1;
      const int hk = vx % 18, hj = (vx / 18) % 6, hi = vx / 108;
      const int gi = i0 - 1 + hi, gj = j0 - 1 + hj, gk = k0 - 1 + hk;
      const bool ok = tt < 2 * kBrickHalo && gi >= 0 && gi < d.px && gj >= 0 && gj < d.py &&
                      gk >= 0 && gk < d.pz;
      const int64_t nv = ((int64_t)gi * d.py + gj) * d.pz + gk;
      // x1 at half resolution (up1): the nearest-neighbour upsampling read in place
      const int64_t nv1 =
          up1 ? ((int64_t)(gi >> 1) * py1 + (gj >> 1)) * pz1 + (gk >> 1) : nv;
      const int c = 16 * gr + 8 * half;
      hv[u] = ok ? *(const uint4*)(c < c1 ? x1 + nv1 * c1 + c : x2 + nv * c2 + (c - c1))
                 : uint4{0u, 0u, 0u, 0u};
    }
  };
  // weight fragments of the channel tile for input-channel group gr, 14 tap pairs
  auto load_b = [&](int gr, Frag8 (&bf)[14]) {
    const int co = blockIdx.y * 16 + r;
    const int ch = 16 * gr + 8 * (g & 1);
#pragma unroll
    for (int c = 0; c < 14; ++c) {
      const int tap = 2 * c + (g >> 1);
      bf[c].u = tap < 27 ? *(const uint4*)(W + ((int64_t)tap * cout + co) * cin + ch)
                         : uint4{0u, 0u, 0u, 0u};
    }
  };
  f32x4 acc[4];
#pragma unroll
  for (int m = 0; m < 4; ++m) acc[m] = f32x4{0.f, 0.f, 0.f, 0.f};
  uint4 hv[6];
  Frag8 bw[14];
  if (items > 0) {
    load_halo(0, hv);
    load_b(gbeg, bw);  // with one group per brick the weights stay for every brick
  }
  for (int t = 0; t < items; ++t) {
    const int gr = gbeg + t % per;
    if (per > 1 && t > 0) load_b(gr, bw);
    __syncthreads();  // the previous item's fragment reads are done
#pragma unroll
    for (int u = 0; u < 6; ++u) {
      const int tt = threadIdx.x + 256 * u;
      if (tt < 2 * kBrickHalo) sx[(tt & 1) * kBrickHalf + (tt >> 1)] = hv[u];
    }
    __syncthreads();
    if (t + 1 < items) load_halo(t + 1, hv);  // in flight during this item's MFMAs
    // B fragments of tap pair c for the wave's four M tiles (LDS offset of the lane's voxel and
    // tap)
    auto read_a = [&](int c, Frag8 (&a)[4]) {
      const int tap = 2 * c + (g >> 1);
      const int tt = tap < 27 ? tap : 26;  // the padding tap reads a valid slot (times 0)
      const int di = tt / 9, dj = (tt / 3) % 3, dk = tt % 3;
      const int base = (g & 1) * kBrickHalf + ((wi + di) * 6 + wj + dj) * 18 + r + dk;
#pragma unroll
      for (int m = 0; m < 4; ++m) a[m].u = sx[base + ((m >> 1) * 6 + (m & 1)) * 18];
    };
    // the next tap pair's four fragments are read while this pair's MFMAs run (the scheduling
    // groups keep that order: four LDS reads, then four MFMAs)
    Frag8 a0[4], a1[4];
    read_a(0, a0);
    __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
#pragma unroll
    for (int c = 0; c < 14; ++c) {
      Frag8(&cur)[4] = (c & 1) ? a1 : a0;
      if (c + 1 < 14) {
        read_a(c + 1, (c & 1) ? a0 : a1);
        __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
      }
#pragma unroll
      for (int m = 0; m < 4; ++m)
        acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bw[c].v, cur[m].v, acc[m], 0, 0, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
    }
    if ((t + 1) % per) continue;  // more groups of this brick follow
    // the brick's last group: lane (g, r) holds channels co0 .. co0 + 3 of voxel k0 + r of each
    // M tile; accumulators cleared for the next brick
    int i0, j0, k0;
    brick_of(t, i0, j0, k0);
    const int kk = k0 + r;
    auto voxel = [&](int m, bool& ok) {
      const int ip = i0 + wi + (m >> 1), jr = j0 + wj + (m & 1);
      ok = ip < d.px && jr < d.py && kk < d.pz;  // past the grid: nothing stored
      return ((int64_t)ip * d.py + jr) * d.pz + kk;
    };
    if (part) {  // split: this split's FP32 partial sums (k_conv_finish adds them)
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        bool ok;
        const int64_t v = voxel(m, ok);
        if (ok) *(f32x4*)(part + ((int64_t)blockIdx.z * d.nv + v) * cout + co0) = acc[m];
        acc[m] = f32x4{0.f, 0.f, 0.f, 0.f};
      }
      continue;
    }
    uint2 o[4];  // the stored BF16 outputs, per M tile
    bool ok0 = false;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      bool ok;
      const int64_t v = voxel(m, ok);
      if (m == 0) ok0 = ok;
      uint16_t* dst =
          co0 < ysplit ? y1 + v * ysplit + co0 : y2 + v * (cout - ysplit) + (co0 - ysplit);
      float val[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) val[e] = acc[m][e] + bco[e];
      if (accum && ok) {
        const uint2 old = *(const uint2*)dst;
        val[0] += bf_lo(old.x);
        val[1] += bf_hi(old.x);
        val[2] += bf_lo(old.y);
        val[3] += bf_hi(old.y);
      }
      if (relu) {
#pragma unroll
        for (int e = 0; e < 4; ++e) val[e] = fmaxf(val[e], 0.f);
      }
      o[m] = uint2{bf_pack(val[0], val[1]), bf_pack(val[2], val[3])};
      if (ok) *(uint2*)dst = o[m];
      acc[m] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    if (fz.pool) {
      // window of the even lane r (k even) and lane r + 1 (k odd), summed in (di, dj, dk) order,
      // dk fastest
      uint2 nb[4];
#pragma unroll
      for (int m = 0; m < 4; ++m) {  // quad_perm [1, 0, 3, 2]: the partner lane's values
        nb[m].x = (uint32_t)__builtin_amdgcn_mov_dpp((int)o[m].x, 0xB1, 0xF, 0xF, false);
        nb[m].y = (uint32_t)__builtin_amdgcn_mov_dpp((int)o[m].y, 0xB1, 0xF, 0xF, false);
      }
      if (!(r & 1) && ok0) {
        float sp[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          sp[0] += bf_lo(o[m].x);
          sp[0] += bf_lo(nb[m].x);
          sp[1] += bf_hi(o[m].x);
          sp[1] += bf_hi(nb[m].x);
          sp[2] += bf_lo(o[m].y);
          sp[2] += bf_lo(nb[m].y);
          sp[3] += bf_hi(o[m].y);
          sp[3] += bf_hi(nb[m].y);
        }
        const int64_t pv = ((int64_t)((i0 + wi) >> 1) * (d.py >> 1) + ((j0 + wj) >> 1)) *
                               (d.pz >> 1) + (kk >> 1);
        *(uint2*)(fz.pool + pv * cout + co0) =
            uint2{bf_pack(0.125f * sp[0], 0.125f * sp[1]), bf_pack(0.125f * sp[2], 0.125f * sp[3])};
      }
    }
    if (fz.yh) {
      // head: A = W (lane (g, r): row r, columns 4 g ..), B = the outputs (lane (g, r): channels
      // 4 g .. of voxel r); D lane (g, r): head channels 4 g + e of voxel r
#pragma unroll
      for (int m = 0; m < 4; ++m) {
        bool ok;
        const int64_t v = voxel(m, ok);
        union {
          uint2 u;
          s16x4 s;
        } b;
        b.u = o[m];
        const f32x4 yv =
            __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(wh, b.s, f32x4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
        if (ok)
          *(uint2*)(fz.yh + v * 16 + 4 * g) =
              uint2{bf_pack(yv[0] + bh[0], yv[1] + bh[1]), bf_pack(yv[2] + bh[2], yv[3] + bh[3])};
      }
    }
  }
}

// y = act(bias + sum over the splits of part[z][v][co]) (+ the existing value with accum), in
// split order; one thread per (voxel, channel).
__global__ void k_conv_finish(const float* __restrict__ part, int ks, VoxDims d, int cout,
                              const float* __restrict__ bias, uint16_t* __restrict__ y1,
                              uint16_t* __restrict__ y2, int ysplit, int relu, int accum) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= d.nv * cout) return;
  const int co = (int)(q % cout);
  const int64_t v = q / cout;
  // up to 8 splits with every load issued at once, then any further splits (in split order)
  float pz[8];
#pragma unroll
  for (int z = 0; z < 8; ++z) pz[z] = z < ks ? part[(int64_t)z * d.nv * cout + q] : 0.f;
  float val = 0.f;
#pragma unroll
  for (int z = 0; z < 8; ++z)
    if (z < ks) val += pz[z];
  for (int z = 8; z < ks; ++z) val += part[(int64_t)z * d.nv * cout + q];
  if (bias) val += bias[co];
  uint16_t* dst = co < ysplit ? y1 + v * ysplit + co : y2 + v * (cout - ysplit) + (co - ysplit);
  if (accum) val += bf2f(*dst);
  if (relu) val = fmaxf(val, 0.f);
  *dst = f2bf(val);
}

// k_conv_finish for a layer whose 2x2x2 average pool is also wanted (ysplit = cout, no accum;
// dimensions even): one thread per (pooled voxel, channel) finishes the window's eight outputs
// as k_conv_finish does, stores them, and stores their pool (the BF16 values summed in FP32 in
// (di, dj, dk) order, dk fastest, times 1/8, rounded to BF16).
__global__ void k_conv_finish_pool(const float* __restrict__ part, int ks, VoxDims d, int cout,
                                   const float* __restrict__ bias, uint16_t* __restrict__ y,
                                   int relu, uint16_t* __restrict__ pool) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int pxc = d.px >> 1, pyc = d.py >> 1, pzc = d.pz >> 1;
  if (q >= (int64_t)pxc * pyc * pzc * cout) return;
  const int co = (int)(q % cout);
  const int64_t pv = q / cout;
  const int pk = (int)(pv % pzc), pj = (int)((pv / pzc) % pyc), pi = (int)(pv / ((int64_t)pzc * pyc));
  int64_t v[8];
  float val[8];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    v[a] = ((int64_t)(2 * pi + (a >> 2)) * d.py + (2 * pj + ((a >> 1) & 1))) * d.pz +
           (2 * pk + (a & 1));
    val[a] = 0.f;
  }
  // up to 8 splits with every load issued at once, then any further splits (in split order)
  float pz[8][8];
#pragma unroll
  for (int z = 0; z < 8; ++z)
#pragma unroll
    for (int a = 0; a < 8; ++a)
      pz[z][a] = z < ks ? part[((int64_t)z * d.nv + v[a]) * cout + co] : 0.f;
#pragma unroll
  for (int z = 0; z < 8; ++z)
#pragma unroll
    for (int a = 0; a < 8; ++a)
      if (z < ks) val[a] += pz[z][a];
  for (int z = 8; z < ks; ++z)
#pragma unroll
    for (int a = 0; a < 8; ++a) val[a] += part[((int64_t)z * d.nv + v[a]) * cout + co];
  float s = 0.f;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    if (bias) val[a] += bias[co];
    if (relu) val[a] = fmaxf(val[a], 0.f);
    const uint16_t o = f2bf(val[a]);
    y[v[a] * cout + co] = o;
    s += bf2f(o);
  }
  pool[pv * cout + co] = f2bf(0.125f * s);
}

// Pointwise 16 -> 16 channel layer (the network head and its data gradient), one thread per
// voxel: y[v][co] = act(bias[co] + sum_ci W[co][ci] x[v][ci]) (+ the existing y with accum), FP32
// sums in ci order; the weights (FP32, staged in LDS) are read as block-wide broadcasts, the
// voxel's 16 channels are two 16-byte loads and stores. Memory-bound: 64 B per voxel.
__global__ __launch_bounds__(256) void k_pointwise16(const uint16_t* __restrict__ x,
                                                     const uint16_t* __restrict__ W,
                                                     const float* __restrict__ bias,
                                                     uint16_t* __restrict__ y, int relu,
                                                     int accum, int64_t nv) {
  __shared__ float4 sw[16 * 4 + 4];  // W[co][ci] as 4 x float4 per co, then the bias
  {
    float* f = (float*)sw;
    f[threadIdx.x] = bf2f(W[threadIdx.x]);
    if (threadIdx.x < 16) f[256 + threadIdx.x] = bias ? bias[threadIdx.x] : 0.f;
  }
  __syncthreads();
  const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  const uint4* xp = (const uint4*)(x + v * 16);
  const uint4 u0 = xp[0], u1 = xp[1];
  const uint32_t q[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
  float xf[16];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    xf[2 * i] = bf2f((uint16_t)(q[i] & 0xFFFFu));
    xf[2 * i + 1] = bf2f((uint16_t)(q[i] >> 16));
  }
  float yo[16];
  if (accum) {
    const uint4* yp = (const uint4*)(y + v * 16);
    const uint4 a0 = yp[0], a1 = yp[1];
    const uint32_t p[8] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      yo[2 * i] = bf2f((uint16_t)(p[i] & 0xFFFFu));
      yo[2 * i + 1] = bf2f((uint16_t)(p[i] >> 16));
    }
  }
  uint32_t o[8];
#pragma unroll
  for (int co = 0; co < 16; ++co) {
    float s = 0.f;
#pragma unroll
    for (int c4 = 0; c4 < 4; ++c4) {
      const float4 w4 = sw[co * 4 + c4];
      s = fmaf(w4.x, xf[4 * c4], s);
      s = fmaf(w4.y, xf[4 * c4 + 1], s);
      s = fmaf(w4.z, xf[4 * c4 + 2], s);
      s = fmaf(w4.w, xf[4 * c4 + 3], s);
    }
    float val = s + ((const float*)sw)[256 + co];
    if (accum) val += yo[co];
    if (relu) val = fmaxf(val, 0.f);
    const uint32_t b = f2bf(val);
    if (co & 1)
      o[co >> 1] |= b << 16;
    else
      o[co >> 1] = b;
  }
  uint4* yq = (uint4*)(y + v * 16);
  yq[0] = uint4{o[0], o[1], o[2], o[3]};
  yq[1] = uint4{o[4], o[5], o[6], o[7]};
}

// ---------------------------------------------------------------------------------------
// Weight gradient on the matrix cores, deterministic (no atomics):
//   dW[tap][co][ci] = sum_v dy[v][co] * x[v + off(tap)][ci],   db[co] = sum_v dy[v][co].
// Per tap a GEMM with M = output channels, N = input channels, K = voxels. Block (x, y) takes
// the (output tile, input tile) pair y = cot * (cin / 16) + cit and walks the 4 x 4 x 16 voxel
// bricks x, x + gridDim.x, ... (the next brick's loads in flight during this one's MFMAs). The
// brick's dy (16 channels) and its input halo (6 x 6 x 18 voxels, 16 channels) are staged in LDS
// transposed to channel-major rows of 16 voxels along z; the halo as three copies shifted by
// dk = -1, 0, 1, so that the 8 consecutive voxels a lane contributes to a tap are one aligned
// 16-byte read. Wave w takes the brick's i-plane w: two K chunks of 32 voxels (lane group g: row
// j = 2 s + g / 2, z-voxels 8 (g % 2) .. + 7), the dy fragments read once for all taps.
// The waves' accumulators are summed through LDS at the end and the block writes its record
// part[x] = {[pair][tap][co % 16][ci % 16] (256 floats per tap and pair), db partials [cout]}
// (bias entries from the input tile 0 blocks). k_wgrad_chunk and k_wgrad_final add the records
// in block order.
// ---------------------------------------------------------------------------------------
constexpr int kWgDS = 256 + 8;            // dy channel row (elements): 528 B, 16 mod 128
constexpr int kWgXS = 6 * 6 * 16 + 8;     // halo channel row (elements): 1168 B, 16 mod 128
constexpr int kWgSmem = 2 * (16 * kWgDS + 3 * 16 * kWgXS);  // 64,512 B
constexpr int kWgChunk = 16;              // records per first-pass sum

template <int TAPS>
__global__ __launch_bounds__(256, 2) void k_conv_wgrad(const uint16_t* __restrict__ x1, int c1,
                                                    const uint16_t* __restrict__ x2, int c2,
                                                    const uint16_t* __restrict__ dy, int cout,
                                                    VoxDims d, float* __restrict__ part) {
  static_assert(2 * TAPS * 256 * 4 + 2 * 16 * 4 <= kWgSmem, "reduction buffer");
  __shared__ __align__(16) unsigned char smem[kWgSmem];
  uint16_t* sD = (uint16_t*)smem;
  uint16_t* sX = sD + 16 * kWgDS;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, r = lane & 15;
  const int cin = c1 + c2, ncit = cin / 16;
  const int npairs = (cout / 16) * ncit;
  const int pr = blockIdx.y, cot = pr / ncit, cit = pr % ncit;
  const int nbk = (d.pz + 15) / 16, nbj = (d.py + 3) / 4;
  const int nbricks = ((d.px + 3) / 4) * nbj * nbk;
  const bool x_first = 16 * cit < c1;  // the input tile lies in x1, else in x2
  const uint16_t* xs = x_first ? x1 : x2;
  const int xc = x_first ? c1 : c2, xoff = x_first ? 16 * cit : 16 * cit - c1;
  auto corner = [&](int b, int& i0, int& j0, int& k0) {
    const int bk = b % nbk, bj = (b / nbk) % nbj, bi = b / (nbk * nbj);
    i0 = 4 * bi;
    j0 = 4 * bj;
    k0 = 16 * bk;
  };
  uint4 dv[2], hv[6];
  // global loads of brick b: dy (thread: 8 channels of one brick voxel, twice) and the halo
  // (8 channels of one halo voxel, up to six times); zero outside the grid
  auto load = [&](int b) {
    int i0, j0, k0;
    corner(b, i0, j0, k0);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int tt = threadIdx.x + 256 * u;
      const int vx = tt & 255, half = tt >> 8;
      const int gi = i0 + (vx >> 6), gj = j0 + ((vx >> 4) & 3), gk = k0 + (vx & 15);
      const bool ok = gi < d.px && gj < d.py && gk < d.pz;
      dv[u] = ok ? *(const uint4*)(dy + (((int64_t)gi * d.py + gj) * d.pz + gk) * cout +
                                   16 * cot + 8 * half)
                 : uint4{0u, 0u, 0u, 0u};
    }
#pragma unroll
    for (int u = 0; u < 6; ++u) {
      const int tt = threadIdx.x + 256 * u;
      const int vx = tt >> 1, half = tt & 1;
      const int hk = vx % 18, hj = (vx / 18) % 6, hi = vx / 108;
      const int gi = i0 - 1 + hi, gj = j0 - 1 + hj, gk = k0 - 1 + hk;
      const bool ok = tt < 2 * 648 && gi >= 0 && gi < d.px && gj >= 0 && gj < d.py && gk >= 0 &&
                      gk < d.pz;
      hv[u] = ok ? *(const uint4*)(xs + (((int64_t)gi * d.py + gj) * d.pz + gk) * xc + xoff +
                                   8 * half)
                 : uint4{0u, 0u, 0u, 0u};
    }
  };
  // the loaded brick into LDS, transposed: sD[co][voxel], sX[copy][ci][hi][hj][m] = halo
  // voxel (hi, hj, m + copy) (copy = dk + 1)
  auto store = [&]() {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int tt = threadIdx.x + 256 * u;
      const int vx = tt & 255, half = tt >> 8;
      const uint32_t q[4] = {dv[u].x, dv[u].y, dv[u].z, dv[u].w};
#pragma unroll
      for (int e = 0; e < 8; ++e)
        sD[(8 * half + e) * kWgDS + vx] = (uint16_t)(q[e >> 1] >> (16 * (e & 1)));
    }
#pragma unroll
    for (int u = 0; u < 6; ++u) {
      const int tt = threadIdx.x + 256 * u;
      if (tt >= 2 * 648) continue;
      const int vx = tt >> 1, half = tt & 1;
      const int hk = vx % 18, row = vx / 18;  // row = hi * 6 + hj
      const uint32_t q[4] = {hv[u].x, hv[u].y, hv[u].z, hv[u].w};
#pragma unroll
      for (int c = (TAPS == 1 ? 1 : 0); c < (TAPS == 1 ? 2 : 3); ++c) {
        const int m = hk - c;
        if (m < 0 || m >= 16) continue;
        uint16_t* dst = sX + (c * 16 + 8 * half) * kWgXS + row * 16 + m;
#pragma unroll
        for (int e = 0; e < 8; ++e) dst[e * kWgXS] = (uint16_t)(q[e >> 1] >> (16 * (e & 1)));
      }
    }
  };
  f32x4 acc[TAPS];
#pragma unroll
  for (int t = 0; t < TAPS; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  float bacc = 0.f;
  const int my = blockIdx.x < nbricks ? (nbricks - 1 - blockIdx.x) / gridDim.x + 1 : 0;
  if (my > 0) load(blockIdx.x);
  for (int t = 0; t < my; ++t) {
    __syncthreads();  // the previous brick's LDS reads are done
    store();
    __syncthreads();
    if (t + 1 < my) load(blockIdx.x + (t + 1) * gridDim.x);
    Frag8 a[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) a[s].u = *(const uint4*)(sD + r * kWgDS + w * 64 + 32 * s + 8 * g);
    if (cit == 0) {
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const uint32_t q[4] = {a[s].u.x, a[s].u.y, a[s].u.z, a[s].u.w};
#pragma unroll
        for (int e = 0; e < 4; ++e)
          bacc += bf2f((uint16_t)(q[e] & 0xFFFFu)) + bf2f((uint16_t)(q[e] >> 16));
      }
    }
#pragma unroll
    for (int tap = 0; tap < TAPS; ++tap) {
      const int tp = TAPS == 1 ? 13 : tap;
      const int di = tp / 9 - 1, dj = (tp / 3) % 3 - 1, dk = tp % 3 - 1;
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        Frag8 b;
        b.u = *(const uint4*)(sX + ((dk + 1) * 16 + r) * kWgXS +
                              ((w + di + 1) * 6 + (2 * s + (g >> 1) + dj + 1)) * 16 + 8 * (g & 1));
        acc[tap] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[s].v, b.v, acc[tap], 0, 0, 0);
      }
    }
  }
  // sum the four waves (lane (r, g) of every wave holds D[co = 4 g + e][ci = r]) and the lane
  // groups' bias sums, in a fixed order
  bacc += __shfl_xor(bacc, 16);
  bacc += __shfl_xor(bacc, 32);
  __syncthreads();
  float* red = (float*)smem;
  float* rb = red + 2 * TAPS * 256;
  if (w >= 2) {
#pragma unroll
    for (int tap = 0; tap < TAPS; ++tap)
#pragma unroll
      for (int e = 0; e < 4; ++e) red[((w - 2) * TAPS + tap) * 256 + e * 64 + lane] = acc[tap][e];
    if (lane < 16) rb[(w - 2) * 16 + lane] = bacc;
  }
  __syncthreads();
  if (w < 2) {
#pragma unroll
    for (int tap = 0; tap < TAPS; ++tap)
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[tap][e] += red[(w * TAPS + tap) * 256 + e * 64 + lane];
    if (lane < 16) bacc += rb[w * 16 + lane];
  }
  __syncthreads();
  if (w == 1) {
#pragma unroll
    for (int tap = 0; tap < TAPS; ++tap)
#pragma unroll
      for (int e = 0; e < 4; ++e) red[tap * 256 + e * 64 + lane] = acc[tap][e];
    if (lane < 16) rb[lane] = bacc;
  }
  __syncthreads();
  if (w != 0) return;
  const int64_t rec = (int64_t)npairs * TAPS * 256 + cout;
  float* P = part + (int64_t)blockIdx.x * rec;
#pragma unroll
  for (int tap = 0; tap < TAPS; ++tap)
#pragma unroll
    for (int e = 0; e < 4; ++e)
      P[((int64_t)pr * TAPS + tap) * 256 + (4 * g + e) * 16 + r] =
          acc[tap][e] + red[tap * 256 + e * 64 + lane];
  if (cit == 0 && lane < 16) P[(int64_t)npairs * TAPS * 256 + 16 * cot + r] = bacc + rb[lane];
}

// First pass of the ordered sum: sum[c][q] = sum of records c * kWgChunk .. + kWgChunk - 1
// (fewer in the last chunk), record length rec.
__global__ void k_wgrad_chunk(const float* __restrict__ part, int nrec, int64_t rec,
                              float* __restrict__ sum) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= rec) return;
  const int b0 = blockIdx.y * kWgChunk, b1 = min(nrec, b0 + kWgChunk);
  float s = 0.f;
  for (int b = b0; b < b1; ++b) s += part[(int64_t)b * rec + q];
  sum[(int64_t)blockIdx.y * rec + q] = s;
}

// Second pass: dW[tap][co][ci] += sum over the chunks (in order), db[co] likewise; dW and db
// adjacent (db = dW + TAPS * cout * cin), one thread per parameter.
template <int TAPS>
__global__ void k_wgrad_final(const float* __restrict__ sum, int nchunk, int cout, int cin,
                              float* __restrict__ dW) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t nw = (int64_t)TAPS * cout * cin;
  if (q >= nw + cout) return;
  const int ncit = cin / 16, npairs = (cout / 16) * ncit;
  const int64_t rec = (int64_t)npairs * TAPS * 256 + cout;
  int64_t idx;
  if (q < nw) {
    const int ci = (int)(q % cin), co = (int)((q / cin) % cout), tap = (int)(q / ((int64_t)cin * cout));
    idx = ((int64_t)((co >> 4) * ncit + (ci >> 4)) * TAPS + tap) * 256 + (co & 15) * 16 + (ci & 15);
  } else {
    idx = (int64_t)npairs * TAPS * 256 + (q - nw);
  }
  float s = 0.f;
  for (int c = 0; c < nchunk; ++c) s += sum[(int64_t)c * rec + idx];
  dW[q] += s;
}

// Wflip[tap'][ci][co] = W[26 - tap'][co][ci] (data gradient = convolution with flipped,
// transposed taps); FP32 master -> BF16.
__global__ void k_w_flip(const float* __restrict__ W, int taps, int cout, int cin,
                         uint16_t* __restrict__ Wf) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t n = (int64_t)taps * cout * cin;
  if (q >= n) return;
  const int ci = (int)(q % cin);
  const int co = (int)((q / cin) % cout);
  const int tap = (int)(q / ((int64_t)cin * cout));
  Wf[((int64_t)(taps - 1 - tap) * cin + ci) * cout + co] = f2bf(W[q]);
}

__global__ void k_f32_to_bf16(const float* __restrict__ a, uint16_t* __restrict__ b, int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n) b[q] = f2bf(a[q]);
}
__global__ void k_bf16_to_f32(const uint16_t* __restrict__ a, float* __restrict__ b, int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n) b[q] = bf2f(a[q]);
}

// dy *= (y > 0)
__global__ void k_relu_bwd(uint16_t* __restrict__ dy, const uint16_t* __restrict__ y,
                           int64_t n) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n && !(bf2f(y[q]) > 0.f)) dy[q] = 0;
}

// Backward of the 2x2x2 average pool (fused into the forward convolutions' epilogues): each fine
// voxel gets 1/8 of its coarse voxel's gradient.
__global__ void k_pool_bwd(const uint16_t* __restrict__ dy, VoxDims dc, int C,
                           uint16_t* __restrict__ dx, VoxDims d, int accum) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= d.nv * C) return;
  const int c = (int)(q % C);
  const int64_t v = q / C;
  int i, j, k;
  vox_decode(v, d, i, j, k);
  const int64_t p = ((int64_t)(i >> 1) * dc.py + (j >> 1)) * dc.pz + (k >> 1);
  float val = 0.125f * bf2f(dy[p * C + c]);
  if (accum) val += bf2f(dx[q]);
  dx[q] = f2bf(val);
}
// nearest-neighbour x2 upsampling (one thread per 8 channels of a fine voxel) and its adjoint
__global__ void k_upsample(const uint16_t* __restrict__ x, VoxDims dc, int C,
                           uint16_t* __restrict__ y, VoxDims d) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int C8 = C / 8;
  if (q >= d.nv * C8) return;
  const int c = 8 * (int)(q % C8);
  int i, j, k;
  const int64_t v = q / C8;
  vox_decode(v, d, i, j, k);
  *(uint4*)(y + v * C + c) =
      *(const uint4*)(x + (((int64_t)(i >> 1) * dc.py + (j >> 1)) * dc.pz + (k >> 1)) * C + c);
}
__global__ void k_upsample_bwd(const uint16_t* __restrict__ dy, VoxDims d, int C,
                               uint16_t* __restrict__ dx, VoxDims dc) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= dc.nv * C) return;
  const int c = (int)(q % C);
  int i, j, k;
  vox_decode(q / C, dc, i, j, k);
  float s = 0.f;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int64_t f = ((int64_t)(2 * i + (a >> 2)) * d.py + (2 * j + ((a >> 1) & 1))) * d.pz +
                      (2 * k + (a & 1));
    s += bf2f(dy[f * C + c]);
  }
  dx[q] = f2bf(s);
}

// Adam on the FP32 master weights; refreshes the BF16 copy used by the forward pass.
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= n) return;
  const float gq = g[q];
  const float mq = b1 * m[q] + (1.f - b1) * gq;
  const float vq = b2 * v[q] + (1.f - b2) * gq * gq;
  m[q] = mq;
  v[q] = vq;
  w[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

}  // namespace amt
