// stream_kernels.hpp — the level-0 kernels of the solvers on the element-column streaming
// operator (stream_op.hpp). Each one has the epilogue of its gather-tile counterpart in
// kernels.hpp / multigrid.hpp (same outputs, same partials); only the operator differs. The
// launch grid is stream_blocks(sg) blocks of kBlock threads for a StreamGrid sg of the same
// (TJ, TK) shape.
#pragma once

#include "multigrid.hpp"
#include "stream_op.hpp"

namespace amt {

struct Raw2 {
  double a, b;
};
struct Raw3 {
  double r, q;
  float w;
};
struct RawRW {
  double r;
  float w;
};
struct Own3f {
  float v[3];
};

// y = K x (+ penalty rows when P != null), x and y of type T.
template <typename T, int TJ, int TK>
__global__ __launch_bounds__(kBlock) void k_apply_s(const T* __restrict__ x, T* __restrict__ y,
                                                    const T* __restrict__ s, HadK<T> K,
                                                    const uint8_t* __restrict__ fixcnt,
                                                    const double* __restrict__ P, GridDims g,
                                                    StreamGrid sg) {
  __shared__ StreamSmem<T, TJ, TK> sm;
  const T pen = P ? (T)(*P) : T(0);
  stream_tile<T, TJ, TK>(
      sm, g, sg, K, s, true, [&](int64_t d) { return x[d]; }, [&](T v) { return v; },
      [&](int64_t n) {
        Own3f f = {{0.f, 0.f, 0.f}};
        if (n >= 0 && P) {
#pragma unroll
          for (int c = 0; c < 3; ++c) f.v[c] = (float)fixcnt[3 * n + c];
        }
        return f;
      },
      [&](int64_t n, T (&yv)[3], const T (&xo)[3], const Own3f& f) {
#pragma unroll
        for (int c = 0; c < 3; ++c) y[3 * n + c] = fma((T)f.v[c] * pen, xo[c], yv[c]);
      });
}

// r = b - K x (with penalty); block partials of |r|^2 (one per block).
template <int TJ, int TK>
__global__ __launch_bounds__(kBlock) void k_residual_s(const double* __restrict__ x,
                                                       const double* __restrict__ b,
                                                       double* __restrict__ r,
                                                       const double* __restrict__ s,
                                                       HadK<double> K,
                                                       const uint8_t* __restrict__ fixcnt,
                                                       const double* __restrict__ P, GridDims g,
                                                       StreamGrid sg, double* partials) {
  __shared__ StreamSmem<double, TJ, TK> sm;
  const double pen = *P;
  double acc[1] = {0.0};
  struct Pre {
    double b[3], f[3];
  };
  stream_tile<double, TJ, TK>(
      sm, g, sg, K, s, true, [&](int64_t d) { return x[d]; }, [&](double v) { return v; },
      [&](int64_t n) {
        Pre o = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
        if (n >= 0) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            o.b[c] = b[3 * n + c];
            o.f[c] = (double)fixcnt[3 * n + c];
          }
        }
        return o;
      },
      [&](int64_t n, double (&yv)[3], const double (&xo)[3], const Pre& o) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double rv = o.b[c] - fma(o.f[c] * pen, xo[c], yv[c]);
          r[3 * n + c] = rv;
          acc[0] = fma(rv, rv, acc[0]);
        }
      });
  store_partials<1>(acc, partials, gridDim.x);
}

// p = z + beta p_old (formed once per node as the plane is stored); q = K p (FP64); partials
// p.q. The flexible-CG operator launch (k_fcg_apply).
template <int TJ, int TK>
__global__ __launch_bounds__(kBlock) void k_fcg_apply_s(
    const double* __restrict__ z, const double* __restrict__ pold, double* __restrict__ pnew,
    double* __restrict__ q, const double* __restrict__ s, HadK<double> K,
    const uint8_t* __restrict__ fixcnt, const double* __restrict__ P, GridDims g, StreamGrid sg,
    const FcgState* st, double* partials) {
  if (st->done) return;
  const double beta = st->beta, pen = *P;
  __shared__ StreamSmem<double, TJ, TK> sm;
  double acc[1] = {0.0};
  stream_tile<double, TJ, TK>(
      sm, g, sg, K, s, true, [&](int64_t d) { return Raw2{z[d], pold[d]}; },
      [&](const Raw2& v) { return fma(beta, v.b, v.a); },
      [&](int64_t n) {
        Own3f f = {{0.f, 0.f, 0.f}};
        if (n >= 0) {
#pragma unroll
          for (int c = 0; c < 3; ++c) f.v[c] = (float)fixcnt[3 * n + c];
        }
        return f;
      },
      [&](int64_t n, double (&yv)[3], const double (&pv)[3], const Own3f& f) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const double qv = fma((double)f.v[c] * pen, pv[c], yv[c]);
          pnew[3 * n + c] = pv[c];
          q[3 * n + c] = qv;
          acc[0] = fma(pv[c], qv, acc[0]);
        }
      });
  store_partials<1>(acc, partials, gridDim.x);
}

// Level-0 pre-smoothing from zero on the FP64 residual r: x0 = wd r, t = r - K x0 (FP32
// operator; clamped DOFs are not restricted). The first V-cycle of a solve (k_mg0_pre).
template <int TJ, int TK>
__global__ __launch_bounds__(kBlock) void k_mg0_pre_s(const double* __restrict__ r, GridDims g,
                                                      StreamGrid sg, const float* __restrict__ s,
                                                      HadK<float> K,
                                                      const uint8_t* __restrict__ fixcnt,
                                                      const double* __restrict__ P,
                                                      const float* __restrict__ wd,
                                                      float* __restrict__ x0,
                                                      float* __restrict__ t, const int* done) {
  if (*done) return;
  __shared__ StreamSmem<float, TJ, TK> sm;
  const float pen = (float)(*P);
  struct Pre {
    float r[3], f[3];
  };
  stream_tile<float, TJ, TK>(
      sm, g, sg, K, s, true, [&](int64_t d) { return RawRW{r[d], wd[d]}; },
      [&](const RawRW& v) { return v.w * (float)v.r; },
      [&](int64_t n) {
        Pre o = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
        if (n >= 0) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            o.r[c] = (float)r[3 * n + c];
            o.f[c] = (float)fixcnt[3 * n + c];
          }
        }
        return o;
      },
      [&](int64_t n, float (&y)[3], const float (&xs)[3], const Pre& o) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int64_t qq = 3 * n + c;
          const float yv = fmaf(o.f[c] * pen, xs[c], y[c]);
          x0[qq] = xs[c];
          t[qq] = o.f[c] > 0.f ? 0.f : o.r[c] - yv;
        }
      });
}

// Level-0 pre-smoothing with the flexible-CG update folded in (k_mg0_pre_upd): alpha = rz / p.q
// from the operator launch's partials; the plane values are wd * (r - alpha q) (r' formed in
// FP64 per node as the plane is stored); the owned nodes write r' to rnew (other blocks still
// read r), x += alpha p and partials of |r'|^2, then x0 and t as the plain pre-smoothing.
template <int TJ, int TK>
__global__ __launch_bounds__(kBlock) void k_mg0_pre_upd_s(
    const double* __restrict__ r, double* __restrict__ rnew, const double* __restrict__ qv,
    double* __restrict__ x, const double* __restrict__ pv, GridDims g, StreamGrid sg,
    const float* __restrict__ s, HadK<float> K, const uint8_t* __restrict__ fixcnt,
    const double* __restrict__ P, const float* __restrict__ wd, float* __restrict__ x0,
    float* __restrict__ t, const FcgState* st, const double* __restrict__ apart, int nba,
    double* __restrict__ upart) {
  double pq[1];
  reduce_partials<1>(apart, nba, nba, pq);
  if (st->done) return;  // uniform over the grid, before any barrier of the tile
  const double a = st->rz / pq[0];
  const float pen = (float)(*P);
  __shared__ StreamSmem<float, TJ, TK> sm;
  double acc[1] = {0.0};
  struct Pre {
    double r[3], q[3], x[3], p[3];
    float f[3];
  };
  stream_tile<float, TJ, TK>(
      sm, g, sg, K, s, true, [&](int64_t d) { return Raw3{r[d], qv[d], wd[d]}; },
      [&](const Raw3& v) { return v.w * (float)fma(-a, v.q, v.r); },
      [&](int64_t n) {
        Pre o;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const bool ok = n >= 0;
          const int64_t qq = ok ? 3 * n + c : 0;
          o.r[c] = ok ? r[qq] : 0.0;
          o.q[c] = ok ? qv[qq] : 0.0;
          o.x[c] = ok ? x[qq] : 0.0;
          o.p[c] = ok ? pv[qq] : 0.0;
          o.f[c] = ok ? (float)fixcnt[qq] : 0.f;
        }
        return o;
      },
      [&](int64_t n, float (&y)[3], const float (&xs)[3], const Pre& o) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int64_t qq = 3 * n + c;
          const double rn = fma(-a, o.q[c], o.r[c]);
          rnew[qq] = rn;
          x[qq] = fma(a, o.p[c], o.x[c]);
          acc[0] = fma(rn, rn, acc[0]);
          const float yv = fmaf(o.f[c] * pen, xs[c], y[c]);
          x0[qq] = xs[c];
          t[qq] = o.f[c] > 0.f ? 0.f : (float)rn - yv;
        }
      });
  store_partials<1>(acc, upart, gridDim.x);
}

// Level-0 post-smoothing into the FP64 preconditioned residual z = x + wd (r - K x), with block
// partials of z.r and z.q (q may be null); k_mg0_post.
template <int TJ, int TK>
__global__ __launch_bounds__(kBlock) void k_mg0_post_s(
    const double* __restrict__ r, GridDims g, StreamGrid sg, const float* __restrict__ s,
    HadK<float> K, const uint8_t* __restrict__ fixcnt, const double* __restrict__ P,
    const float* __restrict__ wd, const float* __restrict__ x, double* __restrict__ z,
    const double* __restrict__ q, double* partials, int stride, const int* done) {
  if (*done) return;
  __shared__ StreamSmem<float, TJ, TK> sm;
  const float pen = (float)(*P);
  double acc[2] = {0.0, 0.0};
  struct Pre {
    double r[3], q[3];
    float w[3], f[3];
  };
  stream_tile<float, TJ, TK>(
      sm, g, sg, K, s, true, [&](int64_t d) { return x[d]; }, [&](float v) { return v; },
      [&](int64_t n) {
        Pre o;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const bool ok = n >= 0;
          const int64_t qq = ok ? 3 * n + c : 0;
          o.r[c] = ok ? r[qq] : 0.0;
          o.q[c] = (ok && q) ? q[qq] : 0.0;
          o.w[c] = ok ? wd[qq] : 0.f;
          o.f[c] = ok ? (float)fixcnt[qq] : 0.f;
        }
        return o;
      },
      [&](int64_t n, float (&y)[3], const float (&xo)[3], const Pre& o) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          const int64_t qq = 3 * n + c;
          const float yv = fmaf(o.f[c] * pen, xo[c], y[c]);
          // wd == 0 marks a DOF no element touches (zero row): its correction stays 0
          const double zv = o.w[c] == 0.f
                                ? 0.0
                                : (double)xo[c] + (double)o.w[c] * (o.r[c] - (double)yv);
          z[qq] = zv;
          acc[0] = fma(zv, o.r[c], acc[0]);
          acc[1] = fma(zv, o.q[c], acc[1]);
        }
      });
  store_partials<2>(acc, partials, stride);
}

}  // namespace amt
