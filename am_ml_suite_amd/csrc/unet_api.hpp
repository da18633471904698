// unet_api.hpp — what the problem code needs of the BF16 U-Net (whose kernels and host code
// are in unet.hpp / unet_net.hpp, compiled in unet.hip): the BF16 conversions, the padded
// voxel grid, the network object and its forward / backward / Adam entry points.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace amt {

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

struct VoxDims {
  int px, py, pz;
  int64_t nv;
};

__device__ __forceinline__ void vox_decode(int64_t v, const VoxDims& d, int& i, int& j, int& k) {
  k = (int)(v % d.pz);
  const int64_t t = v / d.pz;
  j = (int)(t % d.py);
  i = (int)(t / d.py);
}

}  // namespace amt

struct amt_unet {
  struct Layer {
    int taps, c1, c2, cout, relu;
    int64_t w_off, b_off, n_w;
  };
  amt::VoxDims d[4];
  std::vector<Layer> L;
  int64_t n_params = 0;
  float *master = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
  // {weight of the correction in the preconditioner, running numerator, running denominator}
  double* gain = nullptr;
  uint16_t *wb = nullptr, *wflip = nullptr;
  float* part = nullptr;   // partial sums of the split brick convolutions
  float* wpart = nullptr;  // weight-gradient block records and their chunk sums
  int64_t max_w = 0;
  // activations (BF16, channels-last) and their gradients
  enum { X0, A0a, S0, P1, A1a, S1, P2, A2a, S2, P3, B3a, B3, U2, C2a, C2, U1, C1a, C1, U0, C0a,
         C0, OUT, NT_ };
  uint16_t* act[NT_] = {};
  uint16_t* grd[NT_] = {};
  int act_c[NT_] = {}, act_l[NT_] = {};
  int adam_t = 0;
  hipStream_t stream = nullptr;      // stream in use (the attached problem's, or own_stream)
  hipStream_t own_stream = nullptr;
  std::vector<void*> allocs;
  std::string err;
};

namespace amt_unet_impl {
// All three enqueue on net->stream.
void forward(amt_unet* n, bool infer = false);  // act[X0] -> act[OUT] (infer: U* not written)
void backward(amt_unet* n);  // grd[OUT] -> parameter gradients (accumulated into n->grad)
void adam_step(amt_unet* n, float lr);
}  // namespace amt_unet_impl

namespace amt {
// Sets the thread's handle-less error message (amt_last_error(NULL)); returns code.
int set_global_error(int code, const std::string& msg);
}  // namespace amt
