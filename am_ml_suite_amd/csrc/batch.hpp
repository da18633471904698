// batch.hpp — kernels of the batched SIMP engine: many independent problems of different
// sizes advanced in lockstep, each phase ONE persistent launch over all of them.
//
// Work decomposition. Every phase iterates over "items" of one kind: an LDS node tile of the
// operator (BK_TILE), a run of kBElemItem elements (BK_ELEM), of kBDofItem DOFs (BK_DOF) or
// of kBlock nodes (BK_NODE). The items of all problems are numbered consecutively, problem by
// problem, and ip[kind][item] names the owner. A launch has a fixed grid (a multiple of 8
// blocks): the items are cut into 8 contiguous ranges, one per XCD, and the blocks the
// hardware deals to XCD x (blockIdx % 8 == x) stride through range x, so neighbouring tiles of
// one problem share that XCD's L2. An item belongs to exactly one problem, so its reduction
// partial is that problem's, stored at partials[v * stride + item]; a one-block-per-problem
// kernel (or the next phase) sums a problem's partials [item0, item0 + nitem) in fixed order.
//
// Packed layout in HBM: problem p's elements start at e_off, its nodes at n_off (DOFs at
// 3 * n_off); within a problem the single-problem layout (device.hpp) holds.
#pragma once

#include "device.hpp"
#include "stream_op.hpp"

namespace amt {

enum BItemKind { BK_TILE = 0, BK_ELEM, BK_DOF, BK_NODE, BK_KINDS };
constexpr int kBElemItem = 8 * kBlock;  // elements per element item (8 per thread)
constexpr int kBDofItem = 8 * kBlock;   // DOFs per DOF item (8 per thread)

struct BProb {
  GridDims g;
  TileGrid tg;
  int64_t e_off, n_off;  // first element / node of the problem in the packed arrays
  int item0[BK_KINDS], nitem[BK_KINDS];
  int st_off, n_st;      // the problem's filter stencil in the packed stencil table
  int ke;                // the problem's Ke0 in the packed Ke0 tables
  int maxit;
  double penal, Emin, rho_min, move, vf, pf, tol;
};

// Per-problem PCG state (two copies alternate by iteration parity, as in the single-problem
// solver). fresh: set by the start kernel; the first iteration after it does no bookkeeping.
// lastp: parity of the search-direction buffer holding the last p when convergence was seen.
struct BState {
  double rz, rr, rhs2, thresh;
  int it, done, converged, maxit, fresh, lastp;
};

__device__ __forceinline__ int bp_first(int n) {
  const int R = (n + 7) >> 3;
  return (int)(blockIdx.x & 7) * R + (int)(blockIdx.x >> 3);
}
__device__ __forceinline__ int bp_end(int n) {
  const int R = (n + 7) >> 3;
  return min(n, (int)((blockIdx.x & 7) + 1) * R);
}
// Items of this block (all threads of a block see the same sequence).
#define AMT_ITEMS(w, n) for (int w = bp_first(n); w < bp_end(n); w += (int)(gridDim.x >> 3))

// Block total of v stored as the partial of item w.
template <int NV, bool IS_MAX = false>
__device__ __forceinline__ void store_partial_at(const double (&v)[NV], double* partials,
                                                 int stride, int w) {
  double t[NV];
  block_reduce<NV, IS_MAX>(v, t);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) partials[(size_t)q * stride + w] = t[q];
  }
}

// One block per problem: out[p * ostride + q] = sum (or max) of problem p's partials of kind.
template <int NV, bool IS_MAX = false>
__global__ __launch_bounds__(kBlock) void k_b_reduce(const BProb* __restrict__ D, int kind,
                                                     const double* __restrict__ partials,
                                                     int stride, double* __restrict__ out,
                                                     int ostride) {
  const BProb& P = D[blockIdx.x];
  double t[NV];
  reduce_partials<NV, IS_MAX>(partials + P.item0[kind], P.nitem[kind], stride, t);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) out[(size_t)blockIdx.x * ostride + q] = t[q];
  }
}

// SIMP scales (as k_simp, consistent index): s[C-order] from rho[x-fastest slot].
__global__ __launch_bounds__(kBlock) void k_b_simp(const BProb* __restrict__ D,
                                                   const int* __restrict__ ip, int n,
                                                   const double* __restrict__ rho,
                                                   double* __restrict__ s64,
                                                   float* __restrict__ s32) {
  AMT_ITEMS(w, n) {
    const BProb P = D[ip[w]];
    const GridDims g = P.g;
    const int64_t base = (int64_t)(w - P.item0[BK_ELEM]) * kBElemItem;
    for (int l = 0; l < kBElemItem / kBlock; ++l) {
      const int64_t e = base + l * kBlock + threadIdx.x;
      if (e >= g.n_elem) break;
      const int k = (int)(e % g.nz);
      const int64_t t = e / g.nz;
      const int j = (int)(t % g.ny), i = (int)(t / g.ny);
      const double r = rho[P.e_off + ((int64_t)k * g.ny + j) * g.nx + i];
      const double s = (r < P.rho_min) ? 0.0 : P.Emin + pow(r, P.penal) * (1.0 - P.Emin);
      s64[P.e_off + e] = s;
      s32[P.e_off + e] = (float)s;
    }
  }
}

// Diagonal of K without penalty (as k_diag); item partials of max|d|.
__global__ __launch_bounds__(kBlock) void k_b_diag(const BProb* __restrict__ D,
                                                   const int* __restrict__ ip, int n,
                                                   const double* __restrict__ s64,
                                                   const double* __restrict__ Ke64,
                                                   double* __restrict__ diag,
                                                   double* __restrict__ partials) {
  AMT_ITEMS(w, n) {
    const BProb P = D[ip[w]];
    const GridDims g = P.g;
    const double* Ke = Ke64 + (size_t)P.ke * 576;
    const int64_t nd = (int64_t)(w - P.item0[BK_NODE]) * kBlock + threadIdx.x;
    double m[1] = {0.0};
    if (nd < g.n_nodes) {
      int i, j, k;
      node_decode(nd, g, i, j, k);
      double s8[8];
      load_element_scales(s64 + P.e_off, g, i, j, k, s8);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        double d = 0.0;
#pragma unroll
        for (int a = 0; a < 8; ++a) d = fma(s8[a], Ke[(3 * a + c) * 24 + 3 * a + c], d);
        diag[3 * (P.n_off + nd) + c] = d;
        m[0] = fmax(m[0], fabs(d));
      }
    }
    store_partial_at<1, true>(m, partials, n, w);
  }
}

// P = pf * max|K_ii|; Jacobi inverse of the penalised diagonal (zero diagonal -> 1); F_bc;
// item partials of |F_bc|^2. pen[p] receives P.
__global__ __launch_bounds__(kBlock) void k_b_dinv(const BProb* __restrict__ D,
                                                   const int* __restrict__ ip, int n,
                                                   const double* __restrict__ maxdiag,
                                                   const double* __restrict__ diag,
                                                   const uint8_t* __restrict__ fixcnt,
                                                   const double* __restrict__ F,
                                                   double* __restrict__ pen_out,
                                                   double* __restrict__ dinv,
                                                   double* __restrict__ Fb,
                                                   double* __restrict__ partials) {
  AMT_ITEMS(w, n) {
    const int pi = ip[w];
    const BProb P = D[pi];
    const double pen = P.pf * maxdiag[pi];
    if (w == P.item0[BK_DOF] && threadIdx.x == 0) pen_out[pi] = pen;
    const int64_t nd = 3 * P.g.n_nodes, off = 3 * P.n_off;
    const int64_t base = (int64_t)(w - P.item0[BK_DOF]) * kBDofItem;
    double acc[1] = {0.0};
    for (int l = 0; l < kBDofItem / kBlock; ++l) {
      const int64_t i = base + l * kBlock + threadIdx.x;
      if (i >= nd) break;
      double d = diag[off + i];
      const int fc = fixcnt[off + i];
      for (int t = 0; t < fc; ++t) d += pen;  // one K_ii += P per listed occurrence
      dinv[off + i] = (d != 0.0) ? 1.0 / d : 1.0;
      const double fb = fc ? 0.0 : F[off + i];
      Fb[off + i] = fb;
      acc[0] = fma(fb, fb, acc[0]);
    }
    store_partial_at<1>(acc, partials, n, w);
  }
}

// FP64 residual r = F_bc - K u (with penalty), z = (T)(D^-1 r); tile-item partials |r|^2, r.z.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_residual(
    const BProb* __restrict__ D, const int* __restrict__ ip, int n,
    const double* __restrict__ u, const double* __restrict__ Fb, double* __restrict__ r,
    T* __restrict__ z, const double* __restrict__ dinv, const double* __restrict__ s64,
    const double* __restrict__ Ke64, const uint8_t* __restrict__ fixcnt,
    const double* __restrict__ pen, double* __restrict__ partials) {
  __shared__ TileSmem<double> sm;
  AMT_ITEMS(w, n) {
    const int pi = ip[w];
    const BProb P = D[pi];
    const GridDims g = P.g;
    const int64_t off = 3 * P.n_off;
    const int t = w - P.item0[BK_TILE];
    const int i0 = (t / (P.tg.tk * P.tg.tj)) * kTI, j0 = ((t / P.tg.tk) % P.tg.tj) * kTJ,
              k0 = (t % P.tg.tk) * kTK;
    const double* ub = u + off;
    tile_load(sm, g, i0, j0, k0, [&](int64_t q) { return ub[q]; }, s64 + P.e_off);
    __syncthreads();
    int li, lj, lk;
    tile_thread(li, lj, lk);
    const int i = i0 + li, j = j0 + lj, k = k0 + lk;
    double acc[2] = {0.0, 0.0};
    if (i <= g.nx && j <= g.ny && k <= g.nz) {
      const int64_t nd = ((int64_t)i * g.nny + j) * g.nnz + k;
      Nbr<double> nb;
      double s8[8];
      tile_fetch(sm, li, lj, lk, nb, s8);
      double yv[3];
      apply_node(nb, s8, Ke64 + (size_t)P.ke * 576, yv);
      const double pp = pen[pi];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int64_t dof = off + 3 * nd + c;
        yv[c] = fma((double)fixcnt[dof] * pp, nb.v[13][c], yv[c]);
        const double rv = Fb[dof] - yv[c];
        r[dof] = rv;
        const T zv = (T)(dinv[dof] * rv);
        z[dof] = zv;
        acc[0] = fma(rv, rv, acc[0]);
        acc[1] = fma(rv, (double)zv, acc[1]);
      }
    }
    store_partial_at<2>(acc, partials, n, w);  // its barriers also end this item's LDS use
  }
}

// One block per problem: the PCG state from the residual partials (Eigen: rhsNorm2 = |F_bc|^2,
// residualNorm2, absNew). first: iteration count from 0, else carried over from flags (a
// restart from the FP64 residual). Writes S[2p] and flags[p].
__global__ __launch_bounds__(kBlock) void k_b_start(const BProb* __restrict__ D,
                                                    const double* __restrict__ partials,
                                                    int stride, const double* __restrict__ bb,
                                                    int first, BState* __restrict__ S,
                                                    BState* __restrict__ flags) {
  const int pi = blockIdx.x;
  const BProb& P = D[pi];
  double t[2];
  reduce_partials<2>(partials + P.item0[BK_TILE], P.nitem[BK_TILE], stride, t);
  if (threadIdx.x == 0) {
    BState st;
    st.rhs2 = bb[pi];
    st.thresh = fmax(P.tol * P.tol * st.rhs2, 2.2250738585072014e-308);
    st.rr = t[0];
    st.rz = t[1];
    st.it = first ? 0 : flags[pi].it;
    st.maxit = P.maxit;
    st.converged = (st.rhs2 == 0.0) || (t[0] < st.thresh);
    st.done = st.converged || st.it >= st.maxit;
    st.fresh = 1;
    st.lastp = 0;
    S[2 * pi] = st;
    flags[pi] = st;
  }
}

// Bookkeeping at the top of an iteration of parity par (pcg_step_state with the fresh flag
// for t == 0).
__device__ __forceinline__ BState b_step_state(const BState& cur, int par, const double* upart,
                                               int nbu, int stride, double& beta) {
  BState nx = cur;
  beta = 0.0;
  nx.fresh = 0;
  if (cur.done || cur.fresh) return nx;
  double v[2];
  reduce_partials<2>(upart, nbu, stride, v);
  nx.rr = v[0];
  if (!isfinite(v[0]) || !isfinite(v[1])) {  // breakdown: stop, not converged
    nx.done = 1;
    return nx;
  }
  if (v[0] < cur.thresh) {
    nx.done = 1;
    nx.converged = 1;
    nx.lastp = par;  // this iteration would have read its p_old from buffer par
    return nx;
  }
  beta = v[1] / cur.rz;
  nx.rz = v[1];
  nx.it = cur.it + 1;
  if (nx.it >= cur.maxit) nx.done = 1;
  return nx;
}

// Iteration with parity par, first half: p = z + beta p_old (in the LDS tile), q = K p in T
// (FP32 matvec: T = float), tile-item partials p.q.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_apply(
    int par, const BProb* __restrict__ D, const int* __restrict__ ip, int n,
    const T* __restrict__ z, const T* __restrict__ pold, T* __restrict__ pnew,
    T* __restrict__ q, const T* __restrict__ s, const T* __restrict__ Ke,
    const uint8_t* __restrict__ fixcnt, const double* __restrict__ pen, BState* S,
    const double* __restrict__ upart, int ustride, double* __restrict__ apart) {
  __shared__ TileSmem<T> sm;
  AMT_ITEMS(w, n) {
    const int pi = ip[w];
    const BProb P = D[pi];
    const BState cur = S[2 * pi + par];
    double beta_d;
    const BState nx =
        b_step_state(cur, par, upart + P.item0[BK_DOF], P.nitem[BK_DOF], ustride, beta_d);
    if (w == P.item0[BK_TILE] && threadIdx.x == 0) S[2 * pi + (par ^ 1)] = nx;
    if (nx.done) continue;  // uniform over the block
    const GridDims g = P.g;
    const int64_t off = 3 * P.n_off;
    const int t = w - P.item0[BK_TILE];
    const int i0 = (t / (P.tg.tk * P.tg.tj)) * kTI, j0 = ((t / P.tg.tk) % P.tg.tj) * kTJ,
              k0 = (t % P.tg.tk) * kTK;
    const T beta = (T)beta_d;
    const T* zb = z + off;
    const T* pb = pold + off;
    tile_load(sm, g, i0, j0, k0, [&](int64_t qq) { return fma(beta, pb[qq], zb[qq]); },
              s + P.e_off);
    __syncthreads();
    int li, lj, lk;
    tile_thread(li, lj, lk);
    const int i = i0 + li, j = j0 + lj, k = k0 + lk;
    double acc[1] = {0.0};
    if (i <= g.nx && j <= g.ny && k <= g.nz) {
      const int64_t nd = ((int64_t)i * g.nny + j) * g.nnz + k;
      Nbr<T> nb;
      T s8[8];
      tile_fetch(sm, li, lj, lk, nb, s8);
      T yv[3];
      apply_node(nb, s8, Ke + (size_t)P.ke * 576, yv);
      const T pp = (T)pen[pi];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int64_t dof = off + 3 * nd + c;
        yv[c] = fma((T)fixcnt[dof] * pp, nb.v[13][c], yv[c]);
        pnew[dof] = nb.v[13][c];
        q[dof] = yv[c];
        acc[0] = fma((double)nb.v[13][c], (double)yv[c], acc[0]);
      }
    }
    store_partial_at<1>(acc, apart, n, w);
  }
}

// A problem's items of one element-column streaming operator candidate (stream_op.hpp): its
// stream grid (tile shape and plane count of the candidate) and the first of its tiles in the
// candidate's item numbering.
struct BStream {
  StreamGrid sg;
  int item0, nitem;
};

// First half of an iteration on the element-column streaming operator (tile TJ x TK threads,
// each item one stream tile of a problem): the state bookkeeping, outputs and partials of
// k_b_apply, with q = K p from the Hadamard block form HK[P.ke].
template <typename T, int TJ, int TK>
__global__ __launch_bounds__(kBlock) void k_b_apply_s(
    int par, const BProb* __restrict__ D, const BStream* __restrict__ DS,
    const int* __restrict__ ip, int n, const T* __restrict__ z, const T* __restrict__ pold,
    T* __restrict__ pnew, T* __restrict__ q, const T* __restrict__ s,
    const HadK<T>* __restrict__ HK, const uint8_t* __restrict__ fixcnt,
    const double* __restrict__ pen, BState* S, const double* __restrict__ upart, int ustride,
    double* __restrict__ apart) {
  __shared__ StreamSmem<T, TJ, TK> sm;
  struct Raw {
    T a, b;
  };
  struct Own {
    T f[3];
  };
  AMT_ITEMS(w, n) {
    const int pi = ip[w];
    const BProb P = D[pi];
    const BStream BS = DS[pi];
    const BState cur = S[2 * pi + par];
    double beta_d;
    const BState nx =
        b_step_state(cur, par, upart + P.item0[BK_DOF], P.nitem[BK_DOF], ustride, beta_d);
    if (w == BS.item0 && threadIdx.x == 0) S[2 * pi + (par ^ 1)] = nx;
    if (nx.done) continue;  // uniform over the block
    const int64_t off = 3 * P.n_off;
    const T beta = (T)beta_d, pp = (T)pen[pi];
    const T* zb = z + off;
    const T* pb = pold + off;
    const uint8_t* fb = fixcnt + off;
    T* pn = pnew + off;
    T* qb = q + off;
    const HadK<T> K = HK[P.ke];
    double acc[1] = {0.0};
    stream_tile_at<T, TJ, TK>(
        sm, P.g, BS.sg, K, s + P.e_off, w - BS.item0,
        [&](int64_t d) { return Raw{zb[d], pb[d]}; },
        [&](const Raw& v) { return fma(beta, v.b, v.a); },
        [&](int64_t nd) {
          Own o = {{T(0), T(0), T(0)}};
          if (nd >= 0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) o.f[c] = (T)fb[3 * nd + c];
          }
          return o;
        },
        [&](int64_t nd, T (&yv)[3], const T (&pv)[3], const Own& o) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const T qv = fma(o.f[c] * pp, pv[c], yv[c]);
            pn[3 * nd + c] = pv[c];
            qb[3 * nd + c] = qv;
            acc[0] = fma((double)pv[c], (double)qv, acc[0]);
          }
        });
    store_partial_at<1>(acc, apart, n, w);  // its barriers also end this item's LDS use
  }
}

// Second half: alpha = rz / p.q; u += alpha p; r -= alpha q (FP64); z = (T)(D^-1 r); DOF-item
// partials |r|^2, r.z. S[2p + par] is the state the first half wrote; the first half's partials
// of problem p are apart[ar[2p] .. ar[2p] + ar[2p + 1]) (stride astride).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_update(
    int par, const BProb* __restrict__ D, const int* __restrict__ ip, int n,
    double* __restrict__ u, double* __restrict__ r, T* __restrict__ z,
    const T* __restrict__ p, const T* __restrict__ q, const double* __restrict__ dinv,
    const BState* __restrict__ S, const double* __restrict__ apart, const int* __restrict__ ar,
    int astride, double* __restrict__ upart) {
  AMT_ITEMS(w, n) {
    const int pi = ip[w];
    const BProb P = D[pi];
    const BState st = S[2 * pi + par];
    if (st.done) continue;
    double pq[1];
    reduce_partials<1>(apart + ar[2 * pi], ar[2 * pi + 1], astride, pq);
    const double a = st.rz / pq[0];
    const int64_t nd = 3 * P.g.n_nodes, off = 3 * P.n_off;
    const int64_t base = (int64_t)(w - P.item0[BK_DOF]) * kBDofItem;
    double acc[2] = {0.0, 0.0};
    for (int l = 0; l < kBDofItem / kBlock; ++l) {
      const int64_t i = base + l * kBlock + threadIdx.x;
      if (i >= nd) break;
      const int64_t d = off + i;
      u[d] = fma(a, (double)p[d], u[d]);
      const double rv = fma(-a, (double)q[d], r[d]);
      r[d] = rv;
      const T zv = (T)(dinv[d] * rv);
      z[d] = zv;
      acc[0] = fma(rv, rv, acc[0]);
      acc[1] = fma(rv, (double)zv, acc[1]);
    }
    store_partial_at<2>(acc, upart, n, w);
  }
}

// One block: the bookkeeping the next iteration would do, for every problem, into flags.
__global__ __launch_bounds__(kBlock) void k_b_check(int par, const BProb* __restrict__ D,
                                                    int np, const BState* __restrict__ S,
                                                    const double* __restrict__ upart,
                                                    int ustride, BState* __restrict__ flags) {
  for (int pi = 0; pi < np; ++pi) {
    const BProb& P = D[pi];
    double beta;
    const BState nx = b_step_state(S[2 * pi + par], par, upart + P.item0[BK_DOF],
                                   P.nitem[BK_DOF], ustride, beta);
    if (threadIdx.x == 0) flags[pi] = nx;
    __syncthreads();
  }
}

// Residual replacement (van der Vorst and Ye), one block per problem, after every chunk:
// k_b_residual has recomputed r = F_bc - K u in FP64 and z = D^-1 r. par: parity of the state
// slot the chunk left current.
//  - Running problem: the DOF-item partials its next bookkeeping sums are set to the true
//    |r|^2 and r.z (first partial, rest 0), so the next iteration judges convergence on the
//    true residual and forms beta = r.z / rz against its last direction; the state is kept.
//  - Problem stopped as converged on its recursive residual: final if the true one meets the
//    tolerance, else reopened the same way (its last p is carried to buffer 0 when it lies in
//    buffer 1: flags.lastp = 1 asks k_b_carry_p for it).
//  - Problem stopped otherwise (iteration limit, breakdown): stays stopped.
// Writes S[2p + par] and flags[p]; flags.rr is the true |r|^2.
__global__ __launch_bounds__(kBlock) void k_b_replace(int par, const BProb* __restrict__ D,
                                                      const double* __restrict__ tpart,
                                                      int tstride, BState* __restrict__ S,
                                                      double* __restrict__ upart, int ustride,
                                                      BState* __restrict__ flags) {
  const int pi = blockIdx.x;
  const BProb& P = D[pi];
  double t[2];
  reduce_partials<2>(tpart + P.item0[BK_TILE], P.nitem[BK_TILE], tstride, t);
  BState cur = S[2 * pi + par];
  const bool reopen = cur.done && cur.converged && !(t[0] < cur.thresh) && cur.it < cur.maxit;
  const bool run = !cur.done || reopen;
  if (run) {
    for (int b = threadIdx.x; b < P.nitem[BK_DOF]; b += kBlock) {
      double* u0 = upart + P.item0[BK_DOF] + b;
      u0[0] = b == 0 ? t[0] : 0.0;
      u0[ustride] = b == 0 ? t[1] : 0.0;
    }
  }
  if (threadIdx.x == 0) {
    BState f = cur;
    if (reopen) {
      cur.done = 0;
      cur.converged = 0;
      cur.fresh = 0;
      f = cur;             // flags.lastp keeps the buffer of the last p for k_b_carry_p
      cur.lastp = 0;
    } else {
      f.lastp = 0;
    }
    f.rr = t[0];
    S[2 * pi + par] = cur;
    flags[pi] = f;
  }
}

// p_old buffer 0 <- buffer 1 for the problems continuing after a replacement whose last
// direction lies in buffer 1 (the next chunk starts at parity 0).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_b_carry_p(const BProb* __restrict__ D,
                                                      const int* __restrict__ ip, int n,
                                                      const BState* __restrict__ flags,
                                                      T* __restrict__ p0,
                                                      const T* __restrict__ p1) {
  AMT_ITEMS(w, n) {
    const int pi = ip[w];
    const BState f = flags[pi];
    if (f.done || f.lastp != 1) continue;
    const BProb P = D[pi];
    const int64_t nd = 3 * P.g.n_nodes, off = 3 * P.n_off;
    const int64_t base = (int64_t)(w - P.item0[BK_DOF]) * kBDofItem;
    for (int l = 0; l < kBDofItem / kBlock; ++l) {
      const int64_t i = base + l * kBlock + threadIdx.x;
      if (i >= nd) break;
      p0[off + i] = p1[off + i];
    }
  }
}

// DOF-item partials of a.b over each problem's DOFs.
__global__ __launch_bounds__(kBlock) void k_b_dot(const BProb* __restrict__ D,
                                                  const int* __restrict__ ip, int n,
                                                  const double* __restrict__ a,
                                                  const double* __restrict__ b,
                                                  double* __restrict__ partials) {
  AMT_ITEMS(w, n) {
    const BProb P = D[ip[w]];
    const int64_t nd = 3 * P.g.n_nodes, off = 3 * P.n_off;
    const int64_t base = (int64_t)(w - P.item0[BK_DOF]) * kBDofItem;
    double acc[1] = {0.0};
    for (int l = 0; l < kBDofItem / kBlock; ++l) {
      const int64_t i = base + l * kBlock + threadIdx.x;
      if (i >= nd) break;
      acc[0] = fma(a[off + i], b[off + i], acc[0]);
    }
    store_partial_at<1>(acc, partials, n, w);
  }
}

// Sensitivities (as k_sens): per slot e (x-fastest), dc = -p rho^(p-1) u_e^T Ke0 u_e.
__global__ __launch_bounds__(kBlock) void k_b_sens(const BProb* __restrict__ D,
                                                   const int* __restrict__ ip, int n,
                                                   const double* __restrict__ u,
                                                   const double* __restrict__ rho,
                                                   const double* __restrict__ Ke64,
                                                   double* __restrict__ dc) {
  AMT_ITEMS(w, n) {
    const BProb P = D[ip[w]];
    const GridDims g = P.g;
    const double* Ke = Ke64 + (size_t)P.ke * 576;
    const double* ub = u + 3 * P.n_off;
    const int64_t nxy = (int64_t)g.nx * g.ny;
    const int64_t base = (int64_t)(w - P.item0[BK_ELEM]) * kBElemItem;
    for (int l = 0; l < kBElemItem / kBlock; ++l) {
      const int64_t e = base + l * kBlock + threadIdx.x;
      if (e >= g.n_elem) break;
      const int iz = (int)(e / nxy);
      const int64_t rem = e % nxy;
      const int iy = (int)(rem / g.nx), ix = (int)(rem % g.nx);
      double ue[24];
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const int64_t node =
            ((int64_t)(ix + kOX[b]) * g.nny + (iy + kOY[b])) * g.nnz + (iz + kOZ[b]);
        ue[3 * b + 0] = ub[3 * node + 0];
        ue[3 * b + 1] = ub[3 * node + 1];
        ue[3 * b + 2] = ub[3 * node + 2];
      }
      double ce = 0.0;
#pragma unroll
      for (int b = 0; b < 24; ++b) {
        double t = 0.0;
#pragma unroll
        for (int a = 0; a < 24; ++a) t = fma(ue[a], Ke[a * 24 + b], t);
        ce = fma(t, ue[b], ce);
      }
      dc[P.e_off + e] = -P.penal * pow(rho[P.e_off + e], P.penal - 1.0) * ce;
    }
  }
}

// Sensitivity filter (as k_filter) with the problem's own stencil.
__global__ __launch_bounds__(kBlock) void k_b_filter(const BProb* __restrict__ D,
                                                     const int* __restrict__ ip, int n,
                                                     const double* __restrict__ dc,
                                                     const double* __restrict__ rho,
                                                     const StencilEntry* __restrict__ st,
                                                     const double* __restrict__ sw,
                                                     double* __restrict__ dcf) {
#pragma clang fp contract(off)
  AMT_ITEMS(w, n) {
    const BProb P = D[ip[w]];
    const GridDims g = P.g;
    const int64_t nxy = (int64_t)g.nx * g.ny;
    const double* dcb = dc + P.e_off;
    const double* rb = rho + P.e_off;
    const int64_t base = (int64_t)(w - P.item0[BK_ELEM]) * kBElemItem;
    for (int l = 0; l < kBElemItem / kBlock; ++l) {
      const int64_t e = base + l * kBlock + threadIdx.x;
      if (e >= g.n_elem) break;
      const int iz = (int)(e / nxy);
      const int64_t rem = e % nxy;
      const int iy = (int)(rem / g.nx), ix = (int)(rem % g.nx);
      double num = 0.0, den = 0.0;
      for (int t = 0; t < P.n_st; ++t) {
        const StencilEntry o = st[P.st_off + t];
        const int cx = ix + o.dx, cy = iy + o.dy, cz = iz + o.dz;
        if (cx < 0 || cx >= g.nx || cy < 0 || cy >= g.ny || cz < 0 || cz >= g.nz) continue;
        const int64_t m = (int64_t)cz * nxy + (int64_t)cy * g.nx + cx;
        const double wr = sw[P.st_off + t] * rb[m];
        num = num + wr * dcb[m];
        den = den + wr;
      }
      dcf[P.e_off + e] = den > 1e-15 ? num / den : dcb[e];
    }
  }
}

// Design mask from the density (as k_design_mask); element-item partial counts.
__global__ __launch_bounds__(kBlock) void k_b_design(const BProb* __restrict__ D,
                                                     const int* __restrict__ ip, int n,
                                                     const double* __restrict__ rho,
                                                     uint8_t* __restrict__ design,
                                                     double* __restrict__ partials) {
  AMT_ITEMS(w, n) {
    const BProb P = D[ip[w]];
    const int64_t base = (int64_t)(w - P.item0[BK_ELEM]) * kBElemItem;
    double acc[1] = {0.0};
    for (int l = 0; l < kBElemItem / kBlock; ++l) {
      const int64_t e = base + l * kBlock + threadIdx.x;
      if (e >= P.g.n_elem) break;
      const double r = rho[P.e_off + e];
      const uint8_t d = (r > P.rho_min + 1e-6) && (r < 0.99);
      design[P.e_off + e] = d;
      acc[0] += d;
    }
    store_partial_at<1>(acc, partials, n, w);
  }
}

// OC state of every problem at the start of its update (thread per problem).
__global__ void k_b_oc_init(const BProb* __restrict__ D, int np,
                            const double* __restrict__ ndesign, OcState* __restrict__ oc) {
  const int pi = blockIdx.x * blockDim.x + threadIdx.x;
  if (pi >= np) return;
  OcState s;
  s.l1 = 1e-20;
  s.l2 = 1e20;
  s.lmid = 0.0;
  s.vf = D[pi].vf;
  s.n_design = ndesign[pi] > 0 ? ndesign[pi] : (double)D[pi].g.n_elem;
  s.change = s.vol_sum = s.pad = 0.0;
  s.passes = s.done = s.pad1 = s.pad2 = 0;
  oc[pi] = s;
}

// Candidate volumes for the next LEV bisection levels (as k_oc_sweep: the reference's own
// per-element expression), element-item partials.
template <int LEV>
__global__ __launch_bounds__(kBlock) void k_b_oc_sweep(const BProb* __restrict__ D,
                                                       const int* __restrict__ ip, int n,
                                                       const double* __restrict__ rho,
                                                       const double* __restrict__ dcf,
                                                       const uint8_t* __restrict__ design,
                                                       const OcState* __restrict__ oc,
                                                       double* __restrict__ partials) {
  constexpr int NC = (1 << LEV) - 1;
  __shared__ double lmid[NC];
  AMT_ITEMS(w, n) {
    const int pi = ip[w];
    const OcState st = oc[pi];
    if (st.done) continue;
    const BProb P = D[pi];
    __syncthreads();  // the previous item's readers of lmid are done
    if (threadIdx.x < NC) lmid[threadIdx.x] = oc_tree_mid(st.l1, st.l2, threadIdx.x + 1);
    __syncthreads();
    double vol[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) vol[c] = 0.0;
    const int64_t base = (int64_t)(w - P.item0[BK_ELEM]) * kBElemItem;
    for (int l = 0; l < kBElemItem / kBlock; ++l) {
      const int64_t e = base + l * kBlock + threadIdx.x;
      if (e >= P.g.n_elem) break;
      if (!design[P.e_off + e]) continue;
      const double r = rho[P.e_off + e], d = dcf[P.e_off + e];
#pragma unroll
      for (int c = 0; c < NC; ++c) vol[c] += oc_candidate(r, d, lmid[c], P.move, P.rho_min);
    }
    store_partial_at<NC>(vol, partials, n, w);
  }
}

// One block per problem: candidate volumes reduced, bisection tree walked (as k_oc_walk).
template <int LEV>
__global__ __launch_bounds__(kBlock) void k_b_oc_walk(const BProb* __restrict__ D,
                                                      OcState* __restrict__ oc,
                                                      const double* __restrict__ partials,
                                                      int stride) {
  constexpr int NC = (1 << LEV) - 1;
  OcState* st = oc + blockIdx.x;
  if (st->done) return;
  const BProb& P = D[blockIdx.x];
  __shared__ double tots[NC];
  double tot[NC];
  reduce_partials<NC>(partials + P.item0[BK_ELEM], P.nitem[BK_ELEM], stride, tot);
  if (threadIdx.x != 0) return;
#pragma unroll
  for (int c = 0; c < NC; ++c) tots[c] = tot[c];
  double l1 = st->l1, l2 = st->l2;
  int nd = 1, passes = st->passes, done = 0;
  for (int lev = 0; lev < LEV; ++lev) {
    if (passes >= 100 || (l2 - l1) / (l1 + l2) < 1e-4) {
      done = 1;
      break;
    }
    const double lm = 0.5 * (l1 + l2);
    st->lmid = lm;
    ++passes;
    if (tots[nd - 1] / st->n_design > st->vf) {
      l1 = lm;
      nd = 2 * nd + 1;
    } else {
      l2 = lm;
      nd = 2 * nd;
    }
  }
  if (!done && (passes >= 100 || (l2 - l1) / (l1 + l2) < 1e-4)) done = 1;
  st->l1 = l1;
  st->l2 = l2;
  st->passes = passes;
  st->done = done;
}

// rho at the last evaluated lambda (as k_oc_final); element-item partials of max|drho| and of
// the design volume.
__global__ __launch_bounds__(kBlock) void k_b_oc_final(const BProb* __restrict__ D,
                                                       const int* __restrict__ ip, int n,
                                                       double* __restrict__ rho,
                                                       const double* __restrict__ dcf,
                                                       const uint8_t* __restrict__ design,
                                                       const OcState* __restrict__ oc,
                                                       double* __restrict__ partials) {
  AMT_ITEMS(w, n) {
    const int pi = ip[w];
    const BProb P = D[pi];
    const double lm = oc[pi].lmid;
    const int any = oc[pi].passes > 0;
    double acc[1] = {0.0}, mx[1] = {0.0};
    const int64_t base = (int64_t)(w - P.item0[BK_ELEM]) * kBElemItem;
    for (int l = 0; l < kBElemItem / kBlock; ++l) {
      const int64_t e = base + l * kBlock + threadIdx.x;
      if (e >= P.g.n_elem) break;
      const int64_t ge = P.e_off + e;
      const double r = rho[ge];
      double rn = r;
      if (design[ge]) {
        rn = any ? oc_candidate(r, dcf[ge], lm, P.move, P.rho_min) : 0.0;
        acc[0] += rn;
      }
      mx[0] = fmax(mx[0], fabs(rn - r));
      rho[ge] = rn;
    }
    store_partial_at<1, true>(mx, partials, n, w);
    store_partial_at<1>(acc, partials + n, n, w);
  }
}

// One block per problem: change and design volume of the update.
__global__ __launch_bounds__(kBlock) void k_b_oc_final_reduce(const BProb* __restrict__ D,
                                                              OcState* __restrict__ oc,
                                                              const double* __restrict__ partials,
                                                              int stride) {
  const BProb& P = D[blockIdx.x];
  double m[1], v[1];
  reduce_partials<1, true>(partials + P.item0[BK_ELEM], P.nitem[BK_ELEM], stride, m);
  reduce_partials<1>(partials + stride + P.item0[BK_ELEM], P.nitem[BK_ELEM], stride, v);
  if (threadIdx.x == 0) {
    oc[blockIdx.x].change = m[0];
    oc[blockIdx.x].vol_sum = v[0];
  }
}

}  // namespace amt
