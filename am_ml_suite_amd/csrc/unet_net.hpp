// unet_net.hpp — host side of the BF16 3D U-Net: parameters, activations, forward, backward,
// Adam. Included by amt.hip (one translation unit for all device code).
//
// Architecture (4 levels, base 16 channels, 3x3x3 convolutions, ReLU):
//   enc0 16->16->16 | pool | enc1 16->32->32 | pool | enc2 32->64->64 | pool |
//   bottleneck 64->128->128 | up | dec2 (128+64)->64->64 | up | dec1 (64+32)->32->32 | up |
//   dec0 (32+16)->16->16 | head 1x1x1 16->16 (linear; channels 0..2 are the correction)
// Input: 16 channels (0: nodal density feature, 1..3: normalised Jacobi-scaled residual,
// rest zero) on the node grid padded to multiples of 8.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "unet.hpp"


namespace amt_unet_impl {

using amt::VoxDims;

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

inline uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline double normal(uint64_t& s) {
  const double u1 = ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

template <typename T>
int nalloc(amt_unet* n, T** p, size_t count) {
  void* v = nullptr;
  if (hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
    n->err = "hipMalloc failed (U-Net buffers)";
    return AMT_EHIP;
  }
  n->allocs.push_back(v);
  *p = (T*)v;
  return AMT_OK;
}

// Launch shape of a 3x3x3 brick convolution: one channel tile per block (np = 1; more keep too
// many weight fragments live), grid.y = channel tiles, ks splits of the input-channel groups,
// doubled until about 512 blocks exist (the small U-Net levels have few bricks).
struct BrickCfg {
  int np, ny, ks;
  int64_t bricks;
};
inline BrickCfg brick_cfg(const VoxDims& d, int cin, int cout) {
  BrickCfg c;
  c.bricks = (int64_t)((d.px + 3) / 4) * ((d.py + 3) / 4) * ((d.pz + 15) / 16);
  const int nt = cout / 16, ngr = cin / 16;
  c.np = 1;  // one channel tile per block: its 14 weight fragments stay in registers
  c.ny = nt;
  c.ks = 1;
  while (c.ks < ngr && c.bricks * c.ny * c.ks < 512) c.ks *= 2;  // two blocks per CU
  c.ks = std::min(c.ks, ngr);
  const int gpk = (ngr + c.ks - 1) / c.ks;  // groups per split; no split left empty
  c.ks = (ngr + gpk - 1) / gpk;
  return c;
}
// FP32 partial sums the split launches need (floats).
inline int64_t brick_part_floats(const VoxDims& d, int cin, int cout) {
  const BrickCfg c = brick_cfg(d, cin, cout);
  return c.ks > 1 ? (int64_t)c.ks * d.nv * cout : 0;
}

// Launch shape of a weight gradient (k_conv_wgrad): grid.y = channel-tile pairs, grid.x blocks
// walking the bricks (about 512 blocks in all); rec floats per block record, nchunk first-pass
// sums of kWgChunk records.
struct WgradCfg {
  int64_t gx, rec, nchunk;
  int npairs;
};
inline WgradCfg wgrad_cfg(const VoxDims& d, int taps, int cin, int cout) {
  WgradCfg c;
  const int64_t bricks = (int64_t)((d.px + 3) / 4) * ((d.py + 3) / 4) * ((d.pz + 15) / 16);
  c.npairs = (cout / 16) * (cin / 16);
  c.gx = std::max<int64_t>(1, std::min<int64_t>(bricks, 512 / c.npairs));
  c.rec = (int64_t)c.npairs * taps * 256 + cout;
  c.nchunk = ceil_div(c.gx, (int64_t)amt::kWgChunk);
  return c;
}

inline void net_destroy(amt_unet* n) {
  if (!n) return;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  if (n->own_stream) (void)hipStreamSynchronize(n->own_stream);
  for (void* a : n->allocs) (void)hipFree(a);
  if (n->own_stream) (void)hipStreamDestroy(n->own_stream);
  delete n;
}

// px, py, pz: padded grid (multiples of 8).
inline int net_create(int px, int py, int pz, uint64_t seed, amt_unet** out) {
  if (px <= 0 || py <= 0 || pz <= 0 || px % 8 || py % 8 || pz % 8) return AMT_EINVAL;
  amt_unet* n = new amt_unet();
  for (int l = 0; l < 4; ++l) {
    n->d[l].px = px >> l;
    n->d[l].py = py >> l;
    n->d[l].pz = pz >> l;
    n->d[l].nv = (int64_t)n->d[l].px * n->d[l].py * n->d[l].pz;
  }
  struct Spec {
    int taps, c1, c2, cout, relu;
  };
  const Spec specs[15] = {{27, 16, 0, 16, 1},   {27, 16, 0, 16, 1},  {27, 16, 0, 32, 1},
                          {27, 32, 0, 32, 1},   {27, 32, 0, 64, 1},  {27, 64, 0, 64, 1},
                          {27, 64, 0, 128, 1},  {27, 128, 0, 128, 1}, {27, 128, 64, 64, 1},
                          {27, 64, 0, 64, 1},   {27, 64, 32, 32, 1}, {27, 32, 0, 32, 1},
                          {27, 32, 16, 16, 1},  {27, 16, 0, 16, 1},  {1, 16, 0, 16, 0}};
  int64_t off = 0;
  for (const auto& s : specs) {
    amt_unet::Layer L;
    L.taps = s.taps;
    L.c1 = s.c1;
    L.c2 = s.c2;
    L.cout = s.cout;
    L.relu = s.relu;
    L.n_w = (int64_t)s.taps * s.cout * (s.c1 + s.c2);
    L.w_off = off;
    off += L.n_w;
    L.b_off = off;
    off += s.cout;
    n->max_w = std::max(n->max_w, L.n_w);
    n->L.push_back(L);
  }
  n->n_params = off;
  const int chans[amt_unet::NT_] = {16, 16, 16, 16, 32, 32, 32, 64, 64, 64, 128,
                                    128, 128, 64, 64, 64, 32, 32, 32, 16, 16, 16};
  const int lev[amt_unet::NT_] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 2, 1, 1, 1, 0, 0, 0, 0};
  auto bail = [&](int rc) {
    net_destroy(n);
    return rc;
  };
  if (hipStreamCreateWithFlags(&n->own_stream, hipStreamNonBlocking) != hipSuccess)
    return bail(AMT_EHIP);
  n->stream = n->own_stream;
  for (int t = 0; t < amt_unet::NT_; ++t) {
    n->act_c[t] = chans[t];
    n->act_l[t] = lev[t];
    const size_t cnt = (size_t)n->d[lev[t]].nv * chans[t];
    if (nalloc(n, &n->act[t], cnt) || nalloc(n, &n->grd[t], cnt)) return bail(AMT_EHIP);
    if (hipMemset(n->act[t], 0, cnt * 2) != hipSuccess) return bail(AMT_EHIP);
  }
  // partial sums of the split brick convolutions: the largest over every layer, forward and
  // data gradient (output level of each layer in forward order)
  {
    const int out_level[15] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0};
    int64_t need = 0;
    for (size_t li = 0; li < n->L.size(); ++li) {
      const auto& L = n->L[li];
      if (L.taps != 27) continue;
      const VoxDims& dl = n->d[out_level[li]];
      need = std::max(need, brick_part_floats(dl, L.c1 + L.c2, L.cout));
      need = std::max(need, brick_part_floats(dl, L.cout, L.c1 + L.c2));
    }
    if (need > 0 && nalloc(n, &n->part, (size_t)need)) return bail(AMT_EHIP);
    int64_t wneed = 0;
    for (size_t li = 0; li < n->L.size(); ++li) {
      const auto& L = n->L[li];
      const WgradCfg c = wgrad_cfg(n->d[out_level[li]], L.taps, L.c1 + L.c2, L.cout);
      wneed = std::max(wneed, (c.gx + c.nchunk) * c.rec);
    }
    if (nalloc(n, &n->wpart, (size_t)wneed)) return bail(AMT_EHIP);
  }
  if (nalloc(n, &n->master, n->n_params) || nalloc(n, &n->grad, n->n_params) ||
      nalloc(n, &n->m, n->n_params) || nalloc(n, &n->v, n->n_params) ||
      nalloc(n, &n->wb, n->n_params) || nalloc(n, &n->wflip, n->max_w) ||
      nalloc(n, &n->gain, 3) || hipMemset(n->gain, 0, 3 * sizeof(double)) != hipSuccess)
    return bail(AMT_EHIP);
  // He initialisation (ReLU layers), unit-variance head; the gain starts at 0, so an untrained
  // network leaves the preconditioner plain Jacobi
  std::vector<float> h(n->n_params, 0.f);
  uint64_t s = seed ? seed : 0x1234567ull;
  for (size_t li = 0; li < n->L.size(); ++li) {
    const auto& L = n->L[li];
    const bool head = li + 1 == n->L.size();
    const double sd = std::sqrt((head ? 1.0 : 2.0) / (L.taps * (L.c1 + L.c2)));
    for (int64_t q = 0; q < L.n_w; ++q) h[L.w_off + q] = (float)(sd * normal(s));
  }
  if (hipMemcpy(n->master, h.data(), h.size() * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(n->m, 0, h.size() * 4) != hipSuccess ||
      hipMemset(n->v, 0, h.size() * 4) != hipSuccess)
    return bail(AMT_EHIP);
  hipLaunchKernelGGL(amt::k_f32_to_bf16, dim3(ceil_div(n->n_params, 256)), dim3(256), 0,
                     n->stream, n->master, n->wb, n->n_params);
  if (hipStreamSynchronize(n->stream) != hipSuccess) return bail(AMT_EHIP);
  *out = n;
  return AMT_OK;
}

inline void refresh_bf16(amt_unet* n) {
  hipLaunchKernelGGL(amt::k_f32_to_bf16, dim3(ceil_div(n->n_params, 256)), dim3(256), 0,
                     n->stream, n->master, n->wb, n->n_params);
}

// y = conv(x1 ++ x2) with the layer's weights; nt = output channels / 16. 3x3x3 layers run as
// brick convolutions (k_conv_brick, split as brick_cfg says, partial sums in n->part added by
// k_conv_finish); the pointwise head (16 -> 16 channels, one input tensor, one output tensor) as
// k_pointwise16. fz: outputs fused into the layer (its pool, or the head applied to it), in the
// brick epilogue when the launch is unsplit, else (the pool) in k_conv_finish_pool; the caller
// passes a pool only with ysplit = cout and no accum, a head only with cout = 16.
inline void conv_launch(amt_unet* n, int taps, int nt, const uint16_t* x1, int c1,
                        const uint16_t* x2, int c2, const uint16_t* W, const float* bias,
                        uint16_t* y1, uint16_t* y2, int ysplit, int relu, int accum,
                        const VoxDims& d, int up1 = 0, amt::ConvFuse fz = amt::ConvFuse{}) {
  hipStream_t st = n->stream;
  const int cout = 16 * nt;
  if (taps == 27) {
    const BrickCfg c = brick_cfg(d, c1 + c2, cout);
    float* part = c.ks > 1 ? n->part : nullptr;
    // two resident blocks per CU (the kernel's occupancy) walk the bricks, one brick at least
    const int64_t gx = std::max<int64_t>(1, std::min<int64_t>(c.bricks, 512 / (c.ny * c.ks)));
    const dim3 grid((unsigned)gx, (unsigned)c.ny, (unsigned)c.ks), blk(256);
    hipLaunchKernelGGL(amt::k_conv_brick, grid, blk, 0, st, x1, c1, x2, c2, W, bias, y1, y2,
                       ysplit, relu, accum, d, cout, part, up1, part ? amt::ConvFuse{} : fz);
    if (!part) return;
    if (fz.pool) {
      const int64_t cnt = (d.nv / 8) * cout;
      hipLaunchKernelGGL(amt::k_conv_finish_pool, dim3((unsigned)ceil_div(cnt, 256)), dim3(256),
                         0, st, part, c.ks, d, cout, bias, y1, relu, fz.pool);
    } else {
      hipLaunchKernelGGL(amt::k_conv_finish, dim3((unsigned)ceil_div(d.nv * cout, 256)),
                         dim3(256), 0, st, part, c.ks, d, cout, bias, y1, y2, ysplit, relu,
                         accum);
    }
    return;
  }
  // the pointwise head (16 -> 16) and its data gradient; the only 1-tap layer
  (void)x2;
  (void)y2;
  (void)ysplit;
  hipLaunchKernelGGL(amt::k_pointwise16, dim3((unsigned)ceil_div(d.nv, 256)), dim3(256), 0, st,
                     x1, W, bias, y1, relu, accum, d.nv);
}

// up1: in1 is the coarse tensor whose upsampling the layer reads (not written out). pool: the
// tensor that receives the output's 2x2x2 average pool (-1: none). head: the layer after this
// one is the pointwise head, applied in this layer's epilogue (its output written to OUT).
inline void forward_layer(amt_unet* n, int li, int in1, int in2, int out, int up1 = 0,
                          int pool = -1, bool head = false) {
  const auto& L = n->L[li];
  const VoxDims& d = n->d[n->act_l[out]];
  amt::ConvFuse fz{};
  if (pool >= 0) fz.pool = n->act[pool];
  if (head) {
    const auto& H = n->L[li + 1];
    fz.wh = n->wb + H.w_off;
    fz.bh = n->master + H.b_off;
    fz.yh = n->act[amt_unet::OUT];
  }
  conv_launch(n, L.taps, L.cout / 16, n->act[in1], L.c1, in2 >= 0 ? n->act[in2] : nullptr,
              L.c2, n->wb + L.w_off, n->master + L.b_off, n->act[out], nullptr, L.cout, L.relu,
              0, d, up1, fz);
}

inline void upsample(amt_unet* n, int in, int out) {
  const VoxDims& d = n->d[n->act_l[out]];
  const int C = n->act_c[in];
  hipLaunchKernelGGL(amt::k_upsample, dim3(ceil_div(d.nv * C / 8, 256)), dim3(256), 0,
                     n->stream, n->act[in], n->d[n->act_l[in]], C, n->act[out], d);
}

// Forward pass on act[X0] -> act[OUT] (stream-ordered, no sync). With infer, the decoder
// layers read the coarse tensors' upsampling in place and U2, U1, U0 are not written (the
// backward's weight gradient needs them, so training runs with infer = false); the output is
// the same either way.
void forward(amt_unet* n, bool infer) {
  using U = amt_unet;
  forward_layer(n, 0, U::X0, -1, U::A0a);
  forward_layer(n, 1, U::A0a, -1, U::S0, 0, U::P1);
  forward_layer(n, 2, U::P1, -1, U::A1a);
  forward_layer(n, 3, U::A1a, -1, U::S1, 0, U::P2);
  forward_layer(n, 4, U::P2, -1, U::A2a);
  forward_layer(n, 5, U::A2a, -1, U::S2, 0, U::P3);
  forward_layer(n, 6, U::P3, -1, U::B3a);
  forward_layer(n, 7, U::B3a, -1, U::B3);
  if (infer) {
    forward_layer(n, 8, U::B3, U::S2, U::C2a, 1);
  } else {
    upsample(n, U::B3, U::U2);
    forward_layer(n, 8, U::U2, U::S2, U::C2a);
  }
  forward_layer(n, 9, U::C2a, -1, U::C2);
  if (infer) {
    forward_layer(n, 10, U::C2, U::S1, U::C1a, 1);
  } else {
    upsample(n, U::C2, U::U1);
    forward_layer(n, 10, U::U1, U::S1, U::C1a);
  }
  forward_layer(n, 11, U::C1a, -1, U::C1);
  if (infer) {
    forward_layer(n, 12, U::C1, U::S0, U::C0a, 1);
  } else {
    upsample(n, U::C1, U::U0);
    forward_layer(n, 12, U::U0, U::S0, U::C0a);
  }
  forward_layer(n, 13, U::C0a, -1, U::C0, 0, -1, true);  // and the head (layer 14)
}

// Backward of one conv layer. dy = grad of its output (post-ReLU; the ReLU mask is applied
// here). Accumulates dW/db into n->grad; writes the input gradients (dx1, dx2 split) unless
// dx1 < 0; accum_dx adds to existing input gradients.
inline void backward_layer(amt_unet* n, int li, int in1, int in2, int out, int dx1, int dx2,
                           int accum_dx) {
  const auto& L = n->L[li];
  const VoxDims& d = n->d[n->act_l[out]];
  const int64_t cnt = d.nv * L.cout;
  if (L.relu)
    hipLaunchKernelGGL(amt::k_relu_bwd, dim3(ceil_div(cnt, 256)), dim3(256), 0, n->stream,
                       n->grd[out], n->act[out], cnt);
  // weight and bias gradients: block records, then their sums in block order
  const int cin = L.c1 + L.c2;
  const WgradCfg wc = wgrad_cfg(d, L.taps, cin, L.cout);
  const uint16_t* x2 = in2 >= 0 ? n->act[in2] : nullptr;
  float* sums = n->wpart + wc.gx * wc.rec;
  const dim3 wg((unsigned)wc.gx, (unsigned)wc.npairs);
  if (L.taps == 27)
    hipLaunchKernelGGL(amt::k_conv_wgrad<27>, wg, dim3(256), 0, n->stream, n->act[in1], L.c1, x2,
                       L.c2, n->grd[out], L.cout, d, n->wpart);
  else
    hipLaunchKernelGGL(amt::k_conv_wgrad<1>, wg, dim3(256), 0, n->stream, n->act[in1], L.c1, x2,
                       L.c2, n->grd[out], L.cout, d, n->wpart);
  hipLaunchKernelGGL(amt::k_wgrad_chunk, dim3((unsigned)ceil_div(wc.rec, 256), (unsigned)wc.nchunk),
                     dim3(256), 0, n->stream, n->wpart, (int)wc.gx, wc.rec, sums);
  const unsigned gf = (unsigned)ceil_div(L.n_w + L.cout, 256);
  if (L.taps == 27)
    hipLaunchKernelGGL(amt::k_wgrad_final<27>, dim3(gf), dim3(256), 0, n->stream, sums,
                       (int)wc.nchunk, L.cout, cin, n->grad + L.w_off);
  else
    hipLaunchKernelGGL(amt::k_wgrad_final<1>, dim3(gf), dim3(256), 0, n->stream, sums,
                       (int)wc.nchunk, L.cout, cin, n->grad + L.w_off);
  if (dx1 < 0) return;
  hipLaunchKernelGGL(amt::k_w_flip, dim3(ceil_div(L.n_w, 256)), dim3(256), 0, n->stream,
                     n->master + L.w_off, L.taps, L.cout, cin, n->wflip);
  conv_launch(n, L.taps, cin / 16, n->grd[out], L.cout, nullptr, 0, n->wflip, nullptr,
              n->grd[dx1], dx2 >= 0 ? n->grd[dx2] : nullptr, L.c1, 0, accum_dx, d);
}

inline void upsample_bwd(amt_unet* n, int out, int in) {
  const VoxDims& dc = n->d[n->act_l[in]];
  const int C = n->act_c[in];
  hipLaunchKernelGGL(amt::k_upsample_bwd, dim3(ceil_div(dc.nv * C, 256)), dim3(256), 0,
                     n->stream, n->grd[out], n->d[n->act_l[out]], C, n->grd[in], dc);
}
inline void pool_bwd(amt_unet* n, int out, int in, int accum) {
  const VoxDims& d = n->d[n->act_l[in]];
  const int C = n->act_c[in];
  hipLaunchKernelGGL(amt::k_pool_bwd, dim3(ceil_div(d.nv * C, 256)), dim3(256), 0, n->stream,
                     n->grd[out], n->d[n->act_l[out]], C, n->grd[in], d, accum);
}

// Backward from grd[OUT] (after forward). Parameter gradients are accumulated into n->grad
// (zero it first for a fresh gradient).
void backward(amt_unet* n) {
  using U = amt_unet;
  backward_layer(n, 14, U::C0, -1, U::OUT, U::C0, -1, 0);
  backward_layer(n, 13, U::C0a, -1, U::C0, U::C0a, -1, 0);
  backward_layer(n, 12, U::U0, U::S0, U::C0a, U::U0, U::S0, 0);
  upsample_bwd(n, U::U0, U::C1);
  backward_layer(n, 11, U::C1a, -1, U::C1, U::C1a, -1, 0);
  backward_layer(n, 10, U::U1, U::S1, U::C1a, U::U1, U::S1, 0);
  upsample_bwd(n, U::U1, U::C2);
  backward_layer(n, 9, U::C2a, -1, U::C2, U::C2a, -1, 0);
  backward_layer(n, 8, U::U2, U::S2, U::C2a, U::U2, U::S2, 0);
  upsample_bwd(n, U::U2, U::B3);
  backward_layer(n, 7, U::B3a, -1, U::B3, U::B3a, -1, 0);
  backward_layer(n, 6, U::P3, -1, U::B3a, U::P3, -1, 0);
  pool_bwd(n, U::P3, U::S2, 1);
  backward_layer(n, 5, U::A2a, -1, U::S2, U::A2a, -1, 0);
  backward_layer(n, 4, U::P2, -1, U::A2a, U::P2, -1, 0);
  pool_bwd(n, U::P2, U::S1, 1);
  backward_layer(n, 3, U::A1a, -1, U::S1, U::A1a, -1, 0);
  backward_layer(n, 2, U::P1, -1, U::A1a, U::P1, -1, 0);
  pool_bwd(n, U::P1, U::S0, 1);
  backward_layer(n, 1, U::A0a, -1, U::S0, U::A0a, -1, 0);
  backward_layer(n, 0, U::X0, -1, U::A0a, -1, -1, 0);
}

void adam_step(amt_unet* n, float lr) {
  ++n->adam_t;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.f - std::pow(b1, (float)n->adam_t), c2 = 1.f - std::pow(b2, (float)n->adam_t);
  hipLaunchKernelGGL(amt::k_adam, dim3(ceil_div(n->n_params, 256)), dim3(256), 0, n->stream,
                     n->master, n->grad, n->m, n->v, n->n_params, lr, b1, b2, 1e-8f, c1, c2);
  refresh_bf16(n);
}

}  // namespace amt_unet_impl
