// unet.hip — the BF16 3D U-Net as a translation unit of its own: MFMA convolution kernels,
// the network's host code and its handle-level C-ABI (include/amt.h). The problem-side glue
// (sample collection, training loss, preconditioner) is in amt.hip.
#include <algorithm>
#include <cmath>
#include <string>

#include "../../include/amt.h"
#include "device.hpp"
#include "unet_net.hpp"

using namespace amt;

#define AMT_CHECK(call)            \
  do {                             \
    int rc_ = (call);              \
    if (rc_ != AMT_OK) return rc_; \
  } while (0)

namespace {
unsigned nblk(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
}  // namespace

int amt_unet_create(int px, int py, int pz, uint64_t seed, amt_unet** out) {
  if (!out) return amt::set_global_error(AMT_EINVAL, "null argument");
  *out = nullptr;
  const int rc = amt_unet_impl::net_create(px, py, pz, seed, out);
  if (rc != AMT_OK)
    return amt::set_global_error(rc, rc == AMT_EINVAL ? "U-Net grid must be positive multiples of 8"
                                              : "U-Net allocation failed");
  return AMT_OK;
}

int amt_unet_destroy(amt_unet* net) {
  amt_unet_impl::net_destroy(net);
  return AMT_OK;
}

int64_t amt_unet_num_params(const amt_unet* net) { return net ? net->n_params : 0; }

int amt_unet_get_params(amt_unet* net, float* params, int64_t n) {
  if (!net || !params || n != net->n_params) return amt::set_global_error(AMT_EINVAL, "bad buffer");
  if (hipStreamSynchronize(net->stream) != hipSuccess ||
      hipMemcpy(params, net->master, n * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return amt::set_global_error(AMT_EHIP, "copy failed");
  return AMT_OK;
}

int amt_unet_set_params(amt_unet* net, const float* params, int64_t n) {
  if (!net || !params || n != net->n_params) return amt::set_global_error(AMT_EINVAL, "bad buffer");
  if (hipStreamSynchronize(net->stream) != hipSuccess ||
      hipMemcpy(net->master, params, n * 4, hipMemcpyHostToDevice) != hipSuccess)
    return amt::set_global_error(AMT_EHIP, "copy failed");
  amt_unet_impl::refresh_bf16(net);
  net->adam_t = 0;
  if (hipMemsetAsync(net->m, 0, n * 4, net->stream) != hipSuccess ||
      hipMemsetAsync(net->v, 0, n * 4, net->stream) != hipSuccess ||
      hipStreamSynchronize(net->stream) != hipSuccess)
    return amt::set_global_error(AMT_EHIP, "reset failed");
  return AMT_OK;
}

static int net_io(amt_unet* net, int act, const float* in, float* out) {
  const int64_t n = net->d[net->act_l[act]].nv * net->act_c[act];
  float* tmp = nullptr;
  if (hipMalloc(&tmp, n * 4) != hipSuccess) return amt::set_global_error(AMT_EHIP, "hipMalloc failed");
  int rc = AMT_OK;
  if (in) {
    if (hipMemcpyAsync(tmp, in, n * 4, hipMemcpyHostToDevice, net->stream) != hipSuccess)
      rc = AMT_EHIP;
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(n)), dim3(kBlock), 0, net->stream, tmp,
                       act == amt_unet::OUT ? net->grd[act] : net->act[act], n);
  } else {
    hipLaunchKernelGGL(k_bf16_to_f32, dim3(nblk(n)), dim3(kBlock), 0, net->stream,
                       net->act[act], tmp, n);
    if (hipMemcpyAsync(out, tmp, n * 4, hipMemcpyDeviceToHost, net->stream) != hipSuccess)
      rc = AMT_EHIP;
  }
  if (hipStreamSynchronize(net->stream) != hipSuccess) rc = AMT_EHIP;
  (void)hipFree(tmp);
  return rc == AMT_OK ? AMT_OK : amt::set_global_error(rc, "U-Net transfer failed");
}

int amt_unet_benchmark(amt_unet* net, int reps, double* ms_per_forward, double* conv_tflops) {
  if (!net || reps <= 0) return amt::set_global_error(AMT_EINVAL, "invalid argument");
  net->stream = net->own_stream;
  // every convolution's multiply-adds: taps x cin x cout per output voxel of its level
  const int out_level[15] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0};
  double flops = 0.0;
  for (size_t li = 0; li < net->L.size(); ++li) {
    const auto& L = net->L[li];
    flops += 2.0 * L.taps * (L.c1 + L.c2) * L.cout * (double)net->d[out_level[li]].nv;
  }
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return amt::set_global_error(AMT_EHIP, "hipEventCreate");
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return amt::set_global_error(AMT_EHIP, "hipEventCreate");
  }
  amt_unet_impl::forward(net, true);  // warm-up
  (void)hipEventRecord(e0, net->stream);
  for (int r = 0; r < reps; ++r) amt_unet_impl::forward(net, true);
  (void)hipEventRecord(e1, net->stream);
  float ms = 0.f;
  const bool ok = hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!ok) return amt::set_global_error(AMT_EHIP, "forward timing failed");
  const double per = ms / reps;
  if (ms_per_forward) *ms_per_forward = per;
  if (conv_tflops) *conv_tflops = flops / (per * 1e-3) / 1e12;
  return AMT_OK;
}

int amt_unet_infer(amt_unet* net, const float* x, float* y) {
  if (!net || !x || !y) return amt::set_global_error(AMT_EINVAL, "null argument");
  net->stream = net->own_stream;
  AMT_CHECK(net_io(net, amt_unet::X0, x, nullptr));
  amt_unet_impl::forward(net, true);
  return net_io(net, amt_unet::OUT, nullptr, y);
}

int amt_unet_forward(amt_unet* net, const float* x, float* y) {
  if (!net || !x || !y) return amt::set_global_error(AMT_EINVAL, "null argument");
  net->stream = net->own_stream;
  AMT_CHECK(net_io(net, amt_unet::X0, x, nullptr));
  amt_unet_impl::forward(net);
  return net_io(net, amt_unet::OUT, nullptr, y);
}

int amt_unet_activation(amt_unet* net, int tensor, float* out, int64_t n) {
  if (!net || !out) return amt::set_global_error(AMT_EINVAL, "null argument");
  if (tensor < 0 || tensor >= amt_unet::NT_) return amt::set_global_error(AMT_ERANGE, "tensor out of range");
  if (n != net->d[net->act_l[tensor]].nv * net->act_c[tensor])
    return amt::set_global_error(AMT_EINVAL, "size does not match the tensor");
  net->stream = net->own_stream;
  return net_io(net, tensor, nullptr, out);
}

int amt_unet_backward(amt_unet* net, const float* dy, float* dparams) {
  if (!net || !dy || !dparams) return amt::set_global_error(AMT_EINVAL, "null argument");
  net->stream = net->own_stream;
  AMT_CHECK(net_io(net, amt_unet::OUT, dy, nullptr));
  hipLaunchKernelGGL(k_fill<float>, dim3(nblk(net->n_params)), dim3(kBlock), 0, net->stream,
                     net->grad, 0.f, net->n_params);
  amt_unet_impl::backward(net);
  if (hipStreamSynchronize(net->stream) != hipSuccess ||
      hipMemcpy(dparams, net->grad, net->n_params * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return amt::set_global_error(AMT_EHIP, "backward failed");
  return AMT_OK;
}

int amt_unet_get_gain(amt_unet* net, double* gain) {
  if (!net || !gain) return amt::set_global_error(AMT_EINVAL, "null argument");
  if (hipStreamSynchronize(net->stream) != hipSuccess ||
      hipMemcpy(gain, net->gain, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return amt::set_global_error(AMT_EHIP, "gain read failed");
  return AMT_OK;
}

int amt_unet_set_gain(amt_unet* net, double gain) {
  if (!net) return amt::set_global_error(AMT_EINVAL, "null argument");
  if (!std::isfinite(gain)) return amt::set_global_error(AMT_EINVAL, "gain must be finite");
  // restart the running means at this weight
  const double g3[3] = {gain, gain, 1.0};
  if (hipStreamSynchronize(net->stream) != hipSuccess ||
      hipMemcpy(net->gain, g3, sizeof(g3), hipMemcpyHostToDevice) != hipSuccess)
    return amt::set_global_error(AMT_EHIP, "gain write failed");
  return AMT_OK;
}
