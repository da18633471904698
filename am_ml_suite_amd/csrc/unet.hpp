// unet.hpp — BF16 3D U-Net kernels (gfx950): MFMA implicit-GEMM 3x3x3 convolution forward and
// data-gradient, FP32 weight-gradient, pooling/upsampling, ReLU backward, Adam.
//
// Tensors are channels-last on a padded voxel grid (Px, Py, Pz all multiples of 8, z fastest):
// T[((i*Py + j)*Pz + k) * C + c], BF16 (raw 16-bit patterns), channel counts multiples of 16.
// Convolution weights are stored per tap as Wt[tap][co][ci] (taps in dz-fastest order
// (di+1)*9 + (dj+1)*3 + (dk+1)), so the MFMA B operand of 4 consecutive input channels is one
// 8-byte load. Accumulation is FP32.
//
// MFMA layout used (v_mfma_f32_16x16x16_bf16, wave64), checked against a CPU convolution by
// tests/test_gpu_unet.py:
//   A (16x16, M x K): lane l holds A[l % 16][4*(l/16) + e], e = 0..3
//   B (16x16, K x N): lane l holds B[4*(l/16) + e][l % 16]
//   D (16x16, M x N): lane l holds D[4*(l/16) + e][l % 16]
#pragma once

#include "unet_api.hpp"

namespace amt {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------------------
// Convolution (TAPS = 27: 3x3x3 with zero padding; TAPS = 1: pointwise) by MFMA.
//   y[v][co] = act( bias[co] + sum_tap sum_ci W[tap][co][ci] * x[v + off(tap)][ci] )
// Input channels come from x1 (c1 channels) then x2 (c2 channels, may be 0): the U-Net skip
// concatenation is read in place. Output channels [0, ysplit) go to y1 (row stride ysplit),
// [ysplit, cout) to y2 (row stride cout - ysplit) — the data gradient of a concatenated input.
// accum: add to the output instead of overwriting. One wave computes 16 voxels x all output
// channels (NT = cout/16 accumulator tiles); a block of 4 waves covers 64 voxels.
// ---------------------------------------------------------------------------------------
template <int TAPS, int NT>
__global__ __launch_bounds__(256) void k_conv_mfma(
    const uint16_t* __restrict__ x1, int c1, const uint16_t* __restrict__ x2, int c2,
    const uint16_t* __restrict__ W, const float* __restrict__ bias, uint16_t* __restrict__ y1,
    uint16_t* __restrict__ y2, int ysplit, int relu, int accum, VoxDims d) {
  const int lane = threadIdx.x & 63;
  const int64_t v0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * 16;
  if (v0 >= d.nv) return;
  const int cin = c1 + c2;
  const int cout = NT * 16;
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  // A row of this lane: voxel v0 + lane%16
  const int64_t va = v0 + (lane & 15);
  int ai, aj, ak;
  vox_decode(va < d.nv ? va : 0, d, ai, aj, ak);
  const int kg = lane >> 4;  // k-group: channels 4*kg .. 4*kg+3 of the 16-channel chunk
  for (int tap = 0; tap < TAPS; ++tap) {
    const int di = TAPS == 1 ? 0 : tap / 9 - 1, dj = TAPS == 1 ? 0 : (tap / 3) % 3 - 1,
              dk = TAPS == 1 ? 0 : tap % 3 - 1;
    const int ni = ai + di, nj = aj + dj, nk = ak + dk;
    const bool ok = va < d.nv && ni >= 0 && ni < d.px && nj >= 0 && nj < d.py && nk >= 0 &&
                    nk < d.pz;
    const int64_t nvox = ((int64_t)ni * d.py + nj) * d.pz + nk;
    for (int ci0 = 0; ci0 < cin; ci0 += 16) {
      bf16x4 a = {0, 0, 0, 0};
      if (ok) {
        const int ci = ci0 + 4 * kg;
        const uint16_t* src = ci < c1 ? x1 + nvox * c1 + ci : x2 + nvox * c2 + (ci - c1);
        const uint2 raw = *(const uint2*)src;
        a[0] = (short)(raw.x & 0xFFFF);
        a[1] = (short)(raw.x >> 16);
        a[2] = (short)(raw.y & 0xFFFF);
        a[3] = (short)(raw.y >> 16);
      }
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int co = 16 * t + (lane & 15);
        const uint2 raw = *(const uint2*)(W + ((int64_t)tap * cout + co) * cin + ci0 + 4 * kg);
        bf16x4 b;
        b[0] = (short)(raw.x & 0xFFFF);
        b[1] = (short)(raw.x >> 16);
        b[2] = (short)(raw.y & 0xFFFF);
        b[3] = (short)(raw.y >> 16);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
      }
    }
  }
  // D[row = 4*(lane/16) + e][col = lane%16]: row = voxel offset, col = output channel in tile
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int co = 16 * t + (lane & 15);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int64_t v = v0 + 4 * (lane >> 4) + e;
      if (v >= d.nv) continue;
      float val = acc[t][e] + (bias ? bias[co] : 0.f);
      uint16_t* dst = co < ysplit ? y1 + v * ysplit + co : y2 + v * (cout - ysplit) + (co - ysplit);
      if (accum) val += bf2f(*dst);
      if (relu) val = fmaxf(val, 0.f);
      *dst = f2bf(val);
    }
  }
}

// Weight gradient: dW[tap][co][ci] += sum_v dy[v][co] * x[v + off(tap)][ci] over this block's
// voxel range; block (tap, range). Each thread owns (co, ci) pairs; FP32 atomics across ranges.
// Also db[co] += sum_v dy[v][co] (tap 13 blocks only).
template <int TAPS>
__global__ __launch_bounds__(256) void k_conv_wgrad(const uint16_t* __restrict__ x1, int c1,
                                                    const uint16_t* __restrict__ x2, int c2,
                                                    const uint16_t* __restrict__ dy, int cout,
                                                    float* __restrict__ dW,
                                                    float* __restrict__ db, VoxDims d,
                                                    int64_t range) {
  __shared__ float sdy[32][128];
  __shared__ float sx[32][256];
  const int tap = blockIdx.y;
  const int cin = c1 + c2;
  const int di = TAPS == 1 ? 0 : tap / 9 - 1, dj = TAPS == 1 ? 0 : (tap / 3) % 3 - 1,
            dk = TAPS == 1 ? 0 : tap % 3 - 1;
  const int npairs = cout * cin;
  float acc[128];
#pragma unroll
  for (int q = 0; q < 128; ++q) acc[q] = 0.f;
  float bacc = 0.f;
  const int64_t vbeg = (int64_t)blockIdx.x * range;
  const int64_t vend = vbeg + range < d.nv ? vbeg + range : d.nv;
  for (int64_t vc = vbeg; vc < vend; vc += 32) {
    // stage 32 voxels of dy (cout ch) and the shifted x (cin ch)
    for (int q = threadIdx.x; q < 32 * cout; q += 256) {
      const int vv = q / cout, c = q % cout;
      const int64_t v = vc + vv;
      sdy[vv][c] = v < vend ? bf2f(dy[v * cout + c]) : 0.f;
    }
    for (int q = threadIdx.x; q < 32 * cin; q += 256) {
      const int vv = q / cin, c = q % cin;
      const int64_t v = vc + vv;
      float val = 0.f;
      if (v < vend) {
        int i, j, k;
        vox_decode(v, d, i, j, k);
        const int ni = i + di, nj = j + dj, nk = k + dk;
        if (ni >= 0 && ni < d.px && nj >= 0 && nj < d.py && nk >= 0 && nk < d.pz) {
          const int64_t nv = ((int64_t)ni * d.py + nj) * d.pz + nk;
          val = c < c1 ? bf2f(x1[nv * c1 + c]) : bf2f(x2[nv * c2 + (c - c1)]);
        }
      }
      sx[vv][c] = val;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 128; ++q) {
      const int pr = threadIdx.x + 256 * q;
      if (pr < npairs) {
        const int co = pr / cin, ci = pr % cin;
        float s = acc[q];
        for (int vv = 0; vv < 32; ++vv) s = fmaf(sdy[vv][co], sx[vv][ci], s);
        acc[q] = s;
      }
    }
    if (db && tap == TAPS / 2 && threadIdx.x < cout)
      for (int vv = 0; vv < 32; ++vv) bacc += sdy[vv][threadIdx.x];
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < 128; ++q) {
    const int pr = threadIdx.x + 256 * q;
    if (pr < npairs) {
      const int co = pr / cin, ci = pr % cin;
      atomicAdd(dW + ((int64_t)tap * cout + co) * cin + ci, acc[q]);
    }
  }
  if (db && tap == TAPS / 2 && threadIdx.x < cout) atomicAdd(db + threadIdx.x, bacc);
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

// 2x2x2 average pooling (fine dims d, coarse dims d/2), channels C.
__global__ void k_pool(const uint16_t* __restrict__ x, VoxDims d, int C,
                       uint16_t* __restrict__ y, VoxDims dc) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= dc.nv * C) return;
  const int c = (int)(q % C);
  const int64_t v = q / C;
  int i, j, k;
  vox_decode(v, dc, i, j, k);
  float s = 0.f;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int64_t f = ((int64_t)(2 * i + (a >> 2)) * d.py + (2 * j + ((a >> 1) & 1))) * d.pz +
                      (2 * k + (a & 1));
    s += bf2f(x[f * C + c]);
  }
  y[q] = f2bf(0.125f * s);
}
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
// nearest-neighbour x2 upsampling and its adjoint
__global__ void k_upsample(const uint16_t* __restrict__ x, VoxDims dc, int C,
                           uint16_t* __restrict__ y, VoxDims d) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= d.nv * C) return;
  const int c = (int)(q % C);
  int i, j, k;
  vox_decode(q / C, d, i, j, k);
  y[q] = x[(((int64_t)(i >> 1) * dc.py + (j >> 1)) * dc.pz + (k >> 1)) * C + c];
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
