// unet_glue.hpp — kernels between the voxel FE problem and the U-Net tensors.
//
// The network sees the node grid padded to multiples of 8: channel 0 a nodal density feature
// (mean of cbrt(s_e) over the node's elements), channels 1..3 the Jacobi-scaled residual
// D^-1 r divided by its RMS. Its output channels 0..2 are a correction N added to Jacobi:
//   M(r) = D^-1 r + theta * rms * net(rho, D^-1 r / rms)   on free rows,
//   M(r) = D^-1 r                                        on penalty (fixed) rows,
//   M(r) = 0                                             on rows no element touches,
// so M is positively homogeneous and an untrained network (theta = 0) is plain Jacobi.
// Training minimises the energy (U.KU - 2 U.r) / (r.D^-1 r) of U = D^-1 r + theta N, with theta
// set in closed form each step from running means (k_nn_theta), so the scale of the correction
// is right from the first step and the optimiser only shapes N.
#pragma once

#include "kernels.hpp"
#include "unet_api.hpp"

namespace amt {

__device__ __forceinline__ int64_t pad_voxel(const VoxDims& d, int i, int j, int k) {
  return ((int64_t)i * d.py + j) * d.pz + k;
}

// feat[node] = mean over the node's (up to 8) elements of cbrt(s_e)
__global__ void k_nn_feature(const double* __restrict__ s, GridDims g, float* __restrict__ feat) {
  const int64_t n = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.n_nodes) return;
  int i, j, k;
  node_decode(n, g, i, j, k);
  double acc = 0.0;
  int cnt = 0;
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int ei = i - kOX[a], ej = j - kOY[a], ek = k - kOZ[a];
    if (ei >= 0 && ei < g.nx && ej >= 0 && ej < g.ny && ek >= 0 && ek < g.nz) {
      acc += cbrt(s[((int64_t)ei * g.ny + ej) * g.nz + ek]);
      ++cnt;
    }
  }
  feat[n] = cnt ? (float)(acc / cnt) : 0.f;
}

// partials of |D^-1 r|^2 (r FP64 or FP32)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_nn_zz(const T* __restrict__ r,
                                                  const double* __restrict__ dinv, int64_t n,
                                                  double* partials) {
  double acc[1] = {0.0};
  AMT_GRID_STRIDE(i, n) {
    const double z = dinv[i] * (double)r[i];
    acc[0] = fma(z, z, acc[0]);
  }
  store_partials<1>(acc, partials, gridDim.x);
}

// X0 (padded voxels x 16 channels, BF16) from the feature and D^-1 r / rms.
// zz = |D^-1 r|^2 (device scalar); rms = sqrt(zz / n_dof).
template <typename T>
__global__ void k_nn_pack(const T* __restrict__ r, const double* __restrict__ dinv,
                          const float* __restrict__ feat, GridDims g, VoxDims d,
                          const double* __restrict__ zz, uint16_t* __restrict__ X0) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= d.nv * 16) return;
  const int c = (int)(q % 16);
  int i, j, k;
  vox_decode(q / 16, d, i, j, k);
  float val = 0.f;
  if (i <= g.nx && j <= g.ny && k <= g.nz && c < 4) {
    const int64_t n = ((int64_t)i * g.nny + j) * g.nnz + k;
    if (c == 0) {
      val = feat[n];
    } else {
      const double z2 = *zz;
      const double inv = z2 > 0.0 ? sqrt((double)(3 * g.n_nodes) / z2) : 0.0;
      val = (float)(dinv[3 * n + c - 1] * (double)r[3 * n + c - 1] * inv);
    }
  }
  X0[q] = f2bf(val);
}

// z = M(r) (above) with theta the network's gain.
// Writes z in FP64 (preconditioner output) and/or FP32 (training), and partials of z.r, z.q
// (one per block, gridDim.x of them: launch at most the partials region's capacity).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_nn_unpack(
    const uint16_t* __restrict__ OUT, VoxDims d, GridDims g, const T* __restrict__ r,
    const double* __restrict__ dinv, const double* __restrict__ diag,
    const uint8_t* __restrict__ fixcnt, const double* __restrict__ zz,
    const double* __restrict__ theta, double* __restrict__ z64, float* __restrict__ z32,
    const double* __restrict__ q, double* partials) {
  double acc[2] = {0.0, 0.0};
  const double rms = sqrt(*zz / (double)(3 * g.n_nodes));
  const double th = *theta;
  AMT_GRID_STRIDE(t, 3 * g.n_nodes) {
    const int64_t n = t / 3;
    const int c = (int)(t % 3);
    int i, j, k;
    node_decode(n, g, i, j, k);
    const double rv = (double)r[t];
    double zv = 0.0;
    if (fixcnt[t])
      zv = dinv[t] * rv;
    else if (diag[t] != 0.0)
      zv = dinv[t] * rv + th * rms * (double)bf2f(OUT[pad_voxel(d, i, j, k) * 16 + c]);
    if (z64) z64[t] = zv;
    if (z32) z32[t] = (float)zv;
    acc[0] = fma(zv, rv, acc[0]);
    if (q) acc[1] = fma(zv, q[t], acc[1]);
  }
  store_partials<2>(acc, partials, gridDim.x);
}

// Training fields of sample r (FP32): J = D^-1 r on rows M covers, N = rms * OUT[0..2] on the
// free rows the correction reaches (zero elsewhere).
__global__ void k_nn_fields(const uint16_t* __restrict__ OUT, VoxDims d, GridDims g,
                            const float* __restrict__ r, const double* __restrict__ dinv,
                            const double* __restrict__ diag, const uint8_t* __restrict__ fixcnt,
                            const double* __restrict__ zz, float* __restrict__ J,
                            float* __restrict__ N) {
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (t >= 3 * g.n_nodes) return;
  const int64_t n = t / 3;
  const int c = (int)(t % 3);
  int i, j, k;
  node_decode(n, g, i, j, k);
  const bool touched = fixcnt[t] || diag[t] != 0.0;
  const bool corr = !fixcnt[t] && diag[t] != 0.0;
  const double rms = sqrt(*zz / (double)(3 * g.n_nodes));
  J[t] = touched ? (float)(dinv[t] * (double)r[t]) : 0.f;
  N[t] = corr ? (float)(rms * (double)bf2f(OUT[pad_voxel(d, i, j, k) * 16 + c])) : 0.f;
}

// partials of N.r, N.KJ, N.KN, J.r, J.KJ
__global__ __launch_bounds__(kBlock) void k_nn_dots(const float* __restrict__ r,
                                                    const float* __restrict__ J,
                                                    const float* __restrict__ KJ,
                                                    const float* __restrict__ N,
                                                    const float* __restrict__ KN, int64_t n,
                                                    double* partials, int stride) {
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  AMT_GRID_STRIDE(i, n) {
    const double ri = r[i], ji = J[i], kj = KJ[i], ni = N[i], kn = KN[i];
    acc[0] = fma(ni, ri, acc[0]);
    acc[1] = fma(ni, kj, acc[1]);
    acc[2] = fma(ni, kn, acc[2]);
    acc[3] = fma(ji, ri, acc[3]);
    acc[4] = fma(ji, kj, acc[4]);
  }
  store_partials<5>(acc, partials, stride);
}

// The correction's weight is one number for all residuals: theta = E[(N.r - N.KJ) / c] /
// E[N.KN / c] (c = r.D^-1 r), running means over the training steps, i.e. the weight that
// minimises the mean normalised energy of J + theta N. A shared weight makes the sign and
// scale of N consistent across residuals. gain = {theta, E-numerator, E-denominator}; the
// step's theta goes to *theta; terms (optional) gets {U.KU, U.r, c, J.KJ, J.r}.
__global__ void k_nn_theta(const double* __restrict__ dots, const double* __restrict__ rdr,
                           double* __restrict__ gain, double mu, double* __restrict__ theta,
                           double* __restrict__ terms) {
  if (threadIdx.x != 0) return;
  const double nr = dots[0], nkj = dots[1], nkn = dots[2], jr = dots[3], jkj = dots[4];
  const double c = *rdr;
  double th = gain[0];
  if (c > 0.0 && isfinite(nkn) && isfinite(nr) && isfinite(nkj)) {
    const double num = (1.0 - mu) * gain[1] + mu * (nr - nkj) / c;
    const double den = (1.0 - mu) * gain[2] + mu * nkn / c;
    gain[1] = num;
    gain[2] = den;
    if (den > 0.0) th = num / den;
  }
  gain[0] = th;
  *theta = th;
  if (terms) {
    terms[0] = jkj + 2.0 * th * nkj + th * th * nkn;
    terms[1] = jr + th * nr;
    terms[2] = c;
    terms[3] = jkj;
    terms[4] = jr;
  }
}

// Energy-loss gradient w.r.t. the network output: with U = J + theta N and K U = KJ + theta KN,
//   dL/dOUT = rms * theta * 2 (K U - r) / (r.D^-1 r)   on the rows N covers (channels 0..2).
__global__ void k_nn_loss_grad(const float* __restrict__ KJ, const float* __restrict__ KN,
                               const float* __restrict__ r, const double* __restrict__ diag,
                               const uint8_t* __restrict__ fixcnt, GridDims g, VoxDims d,
                               const double* __restrict__ zz, const double* __restrict__ theta,
                               const double* __restrict__ rdr, uint16_t* __restrict__ dOUT) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= d.nv * 16) return;
  const int c = (int)(q % 16);
  int i, j, k;
  vox_decode(q / 16, d, i, j, k);
  float val = 0.f;
  const double th = *theta, cc = *rdr;
  if (i <= g.nx && j <= g.ny && k <= g.nz && c < 3 && cc > 0.0) {
    const int64_t t = 3 * (((int64_t)i * g.nny + j) * g.nnz + k) + c;
    if (!fixcnt[t] && diag[t] != 0.0) {
      const double rms = sqrt(*zz / (double)(3 * g.n_nodes));
      const double ku = (double)KJ[t] + th * (double)KN[t];
      val = (float)(rms * th * 2.0 * (ku - (double)r[t]) / cc);
    }
  }
  dOUT[q] = f2bf(val);
}

// u = gamma z with gamma = (z.F) / (z.Kz) (the energy-optimal multiple of z); dots = {z.F, z.Kz}
__global__ void k_nn_scale(double* __restrict__ u, const double* __restrict__ z,
                           const double* __restrict__ dots, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double den = dots[1];
  const double gamma = (den > 0.0 && isfinite(den) && isfinite(dots[0])) ? dots[0] / den : 0.0;
  u[i] = gamma * z[i];
}

// u += gamma z with gamma = dots[0] / dots[1] (0 unless the denominator is positive and both are
// finite): the energy-optimal step along z for the residual the dots were taken with.
__global__ void k_nn_step(double* __restrict__ u, const double* __restrict__ z,
                          const double* __restrict__ dots, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double den = dots[1];
  const double gamma = (den > 0.0 && isfinite(den) && isfinite(dots[0])) ? dots[0] / den : 0.0;
  u[i] = fma(gamma, z[i], u[i]);
}

// partials of r . D^-1 r
__global__ __launch_bounds__(kBlock) void k_nn_rdr(const float* __restrict__ r,
                                                   const double* __restrict__ dinv, int64_t n,
                                                   double* partials) {
  double acc[1] = {0.0};
  AMT_GRID_STRIDE(i, n) acc[0] = fma((double)r[i] * dinv[i], (double)r[i], acc[0]);
  store_partials<1>(acc, partials, gridDim.x);
}

__global__ void k_d2f(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

__global__ void k_f2d(const float* __restrict__ a, double* __restrict__ b, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}

}  // namespace amt
