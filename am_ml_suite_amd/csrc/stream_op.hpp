// stream_op.hpp — element-column streaming form of the matrix-free hex8 operator.
//
// Element product in the Hadamard basis. Under the reflections x -> -x, y -> -y, z -> -z of
// the cube element, with the displacement component along the reflected axis changing sign,
// the isotropic Ke0 commutes with all three, so in the basis of their joint eigenvectors it is
// block diagonal: 8 blocks of 3x3. With the corner index b = ox + 2 oy + 4 oz, the eigenvector
// for corner mode m and component c is H[b][m] = (-1)^popcount(b & m) in component c, and it
// belongs to class q = m ^ (1 << c). So
//   s Ke0 u = T (s Km) T^T u,   T^T u = three 8-point Walsh-Hadamard transforms (24 add/sub
//   each), Km = T^T Ke0 T / 64 block diagonal,
// and one element costs ~190 operations (72 + 72 add/sub in the transforms, ~45 in the blocks)
// instead of the 1,152 flops of the dense 24 x 24 product. The blocks come from the same Ke0
// on the host (HadK).
//
// Streaming tile. A block of TJ x TK = 256 threads owns a column of element positions
// (j, k) each and walks the element planes i = i0 - 1 .. i0 + L - 1. Per plane it
//   - writes the next node plane (TJ + 1) x (TK + 1) into LDS (two or three plane buffers),
//     from values loaded one step earlier (so the load latency hides behind a step's work);
//   - forms its element's 24 inputs from two node planes and applies s_e Ke0;
//   - adds the element's ox = 0 corners to the previous element's ox = 1 corners, kept in
//     registers: the sum along x never leaves the thread;
//   - keeps the partial sums of its own node (corner (0, 0)) in registers, exchanges the other 9
//     (3 corners (oy, oz) x 3 components) through LDS, and each owned node adds its four in a
//     fixed order.
// A tile owns (TJ - 1) x (TK - 1) node columns and L node planes; neighbouring tiles overlap by
// one element column. Per node this is ~3 LDS values written and ~27 read, instead of the 27
// neighbours x 8 elements of a gather tile, and every node's eight contributions are summed in
// the same order whatever TJ, TK and L are: all tile shapes give bitwise identical operator
// values.
#pragma once

#include <algorithm>
#include <cmath>

namespace amt {

// Km's eight 3x3 blocks (class q, component rows c and columns c'). For the isotropic cube the
// 48 symmetries of the cube fix them up to eight numbers (verified on the host, see
// hadamard_blocks):
//   q = 0 (modes x_0 = (1, x), (2, y), (4, z)):      a on the diagonal, b off it;
//   q = 7 (modes (6, x), (5, y), (3, z)):           g on the diagonal, h off it;
//   q = 1 << z: the translation (mode 0, component z) carries no energy, the other two
//     components pair as [[c, d], [d, c]];
//   q = 7 ^ (1 << w): component w alone with f; the other two pair as e [[1, 1], [1, 1]]
//     (the rotation about axis w carries no energy).
template <typename T>
struct HadK {
  T amb, b, c, d, e, f, gmh, h;  // amb = a - b, gmh = g - h
};

// The eight numbers from Km = T^T Ke T / 64, Ke the 24 x 24 Ke0 in the reference's local node
// order (element_kernel.cpp:196-199). Returns false unless all 576 entries of Km match the
// pattern above to 1e-12 of its largest entry (they do for any isotropic material; rounding
// leaves ~1e-17 where the pattern has zeros). Host code.
template <typename T>
inline bool hadamard_blocks(const double* Ke, HadK<T>& out) {
  static const int ox[8] = {0, 1, 1, 0, 0, 1, 1, 0}, oy[8] = {0, 0, 1, 1, 0, 0, 1, 1},
                   oz[8] = {0, 0, 0, 0, 1, 1, 1, 1};
  int bof[8];
  for (int a = 0; a < 8; ++a) bof[a] = ox[a] + 2 * oy[a] + 4 * oz[a];
  auto H = [&](int a, int m) { return (__builtin_popcount(bof[a] & m) & 1) ? -1.0 : 1.0; };
  double Km[24][24];  // [3 m + c][3 m' + c']
  double mx = 0.0;
  for (int m = 0; m < 8; ++m)
    for (int c = 0; c < 3; ++c)
      for (int m2 = 0; m2 < 8; ++m2)
        for (int c2 = 0; c2 < 3; ++c2) {
          double acc = 0.0;
          for (int a = 0; a < 8; ++a)
            for (int a2 = 0; a2 < 8; ++a2)
              acc += H(a, m) * Ke[(3 * a + c) * 24 + 3 * a2 + c2] * H(a2, m2);
          Km[3 * m + c][3 * m2 + c2] = acc / 64.0;
          mx = std::max(mx, std::fabs(acc / 64.0));
        }
  auto B = [&](int q, int c, int c2) { return Km[3 * (q ^ (1 << c)) + c][3 * (q ^ (1 << c2)) + c2]; };
  const double a = B(0, 0, 0), b = B(0, 0, 1), g = B(7, 0, 0), h = B(7, 0, 1);
  const double c = B(1, 1, 1), d = B(1, 1, 2), e = B(6, 1, 1), f = B(6, 0, 0);
  // the pattern's prediction for every entry of Km
  for (int m = 0; m < 8; ++m)
    for (int cc = 0; cc < 3; ++cc)
      for (int m2 = 0; m2 < 8; ++m2)
        for (int c2 = 0; c2 < 3; ++c2) {
          const int q = m ^ (1 << cc), q2 = m2 ^ (1 << c2);
          double pred = 0.0;
          if (q == q2) {
            if (q == 0 || q == 7) {
              pred = cc == c2 ? (q == 0 ? a : g) : (q == 0 ? b : h);
            } else if (__builtin_popcount(q) == 1) {
              const int z = __builtin_ctz(q);
              pred = (cc == z || c2 == z) ? 0.0 : (cc == c2 ? c : d);
            } else {
              const int w = __builtin_ctz(7 ^ q);
              pred = (cc == w && c2 == w) ? f : ((cc == w || c2 == w) ? 0.0 : e);
            }
          }
          if (std::fabs(Km[3 * m + cc][3 * m2 + c2] - pred) > 1e-12 * mx) return false;
        }
  out.amb = (T)(a - b);
  out.b = (T)b;
  out.c = (T)c;
  out.d = (T)d;
  out.e = (T)e;
  out.f = (T)f;
  out.gmh = (T)(g - h);
  out.h = (T)h;
  return true;
}

template <typename T>
__device__ __forceinline__ void fwht8(T (&v)[8]) {
#pragma unroll
  for (int h = 1; h < 8; h <<= 1)
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (!(i & h)) {
        const T a = v[i], b = v[i + h];
        v[i] = a + b;
        v[i + h] = a - b;
      }
}

// u[c][b] (corner b = ox + 2 oy + 4 oz) <- s Ke0 u, in place: transform, the eight blocks of
// s Km (u[c][m] holds mode m of component c), transform back.
template <typename T>
__device__ __forceinline__ void had_apply(T (&u)[3][8], const HadK<T>& K, T s) {
  fwht8(u[0]);
  fwht8(u[1]);
  fwht8(u[2]);
  const T amb = s * K.amb, bs = s * K.b, cs = s * K.c, ds = s * K.d, es = s * K.e,
          fs = s * K.f, gmh = s * K.gmh, hs = s * K.h;
  T w[3][8];
  {  // q = 0
    const T t = bs * (u[0][1] + u[1][2] + u[2][4]);
    w[0][1] = fma(amb, u[0][1], t);
    w[1][2] = fma(amb, u[1][2], t);
    w[2][4] = fma(amb, u[2][4], t);
  }
  {  // q = 7
    const T t = hs * (u[0][6] + u[1][5] + u[2][3]);
    w[0][6] = fma(gmh, u[0][6], t);
    w[1][5] = fma(gmh, u[1][5], t);
    w[2][3] = fma(gmh, u[2][3], t);
  }
  // q = 1, 2, 4: translations carry nothing; the pairs (3 y, 5 z), (3 x, 6 z), (5 x, 6 y)
  w[0][0] = w[1][0] = w[2][0] = T(0);
  w[1][3] = fma(cs, u[1][3], ds * u[2][5]);
  w[2][5] = fma(cs, u[2][5], ds * u[1][3]);
  w[0][3] = fma(cs, u[0][3], ds * u[2][6]);
  w[2][6] = fma(cs, u[2][6], ds * u[0][3]);
  w[0][5] = fma(cs, u[0][5], ds * u[1][6]);
  w[1][6] = fma(cs, u[1][6], ds * u[0][5]);
  {  // q = 6 (w = x): (7 x) alone; (4 y, 2 z) pair
    const T t = es * (u[1][4] + u[2][2]);
    w[1][4] = t;
    w[2][2] = t;
    w[0][7] = fs * u[0][7];
  }
  {  // q = 5 (w = y): (7 y) alone; (4 x, 1 z) pair
    const T t = es * (u[0][4] + u[2][1]);
    w[0][4] = t;
    w[2][1] = t;
    w[1][7] = fs * u[1][7];
  }
  {  // q = 3 (w = z): (7 z) alone; (2 x, 1 y) pair
    const T t = es * (u[0][2] + u[1][1]);
    w[0][2] = t;
    w[1][1] = t;
    w[2][7] = fs * u[2][7];
  }
  fwht8(w[0]);
  fwht8(w[1]);
  fwht8(w[2]);
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int b = 0; b < 8; ++b) u[c][b] = w[c][b];
}

// Tile decomposition of the node grid: tiles of (TJ - 1) x (TK - 1) node columns and L node
// planes. Blocks map to tiles XCD-contiguously (block b runs on XCD b % 8 and takes tile
// (b % 8) * per_xcd + b / 8), and tiles are numbered with the x segment slowest: an XCD's range
// is whole slabs of column tiles, so the j and k halos a tile shares with its neighbours are read
// through the same L2 (numbering the x segments of one column adjacently instead cut the ranges
// through the columns and fetched those halos twice: 1.15-1.54x of the operator's bytes against
// 1.02-1.07x).
struct StreamGrid {
  int tj, tk, L;  // tile shape (threads along j, k) and node planes per tile
  int nj, nk, ni;
  int n_tiles, per_xcd;
};

inline StreamGrid make_stream_grid(const GridDims& g, int TJ, int TK, int L) {
  StreamGrid s;
  s.tj = TJ;
  s.tk = TK;
  s.L = L;
  s.nj = (g.ny + 1 + TJ - 2) / (TJ - 1);
  s.nk = (g.nz + 1 + TK - 2) / (TK - 1);
  s.ni = (g.nx + 1 + L - 1) / L;
  s.n_tiles = s.nj * s.nk * s.ni;
  s.per_xcd = (s.n_tiles + 7) / 8;
  return s;
}
inline unsigned stream_blocks(const StreamGrid& s) { return 8u * (unsigned)s.per_xcd; }

// Row stride of an LDS node plane: in FP32 a 32-lane group reads two 16-node rows when TK is
// 16, so they are placed 48 words apart (16 banks).
template <typename T, int TK>
constexpr int kPlaneRow = (sizeof(T) == 4 && TK == 16) ? 48 : TK + 1;

// FP32 tiles take one barrier per plane step: three node-plane buffers and two slot buffers,
// each node's sum finished one step later. FP64 tiles keep two barriers per step and two plane
// buffers, so that four blocks of the largest shape fit the LDS of a CU.
template <typename T>
constexpr bool kOneBarrier = sizeof(T) == 4;

template <typename T, int TJ, int TK>
struct StreamSmem {
  // [buffer][component][row * stride + col]
  T plane[kOneBarrier<T> ? 3 : 2][3][(TJ + 1) * kPlaneRow<T, TK>];
  // corners (oy, oz) != (0, 0) of each element column; corner (0, 0) is the thread's own node and
  // stays in registers
  T slot[kOneBarrier<T> ? 2 : 1][9][TJ * TK];  // [buffer][(oy + 2 oz - 1) * 3 + c][thread]
};

// Runs tile t of the grid (t outside [0, n_tiles): nothing). ld(q) issues the loads of DOF q's
// raw inputs and returns them; val(raw) forms the operand value from them (the caller's
// transform, e.g. p = z + beta p_old), applied when the value is stored to LDS one step later.
// pre(n) issues the owned node n's epilogue loads at the start of its step; fin(n, y, xo, pre)
// gets y = K x on the node's three rows (without penalty) and xo = x at the node. All threads
// of the block call it with the same t (barriers).
template <typename T, int TJ, int TK, typename LD, typename VAL, typename PRE, typename FIN>
__device__ __forceinline__ void stream_tile_at(StreamSmem<T, TJ, TK>& sm, const GridDims& g,
                                               const StreamGrid& sg, const HadK<T>& K,
                                               const T* __restrict__ s, int t, LD ld, VAL val,
                                               PRE pre, FIN fin) {
  static_assert(TJ * TK == kBlock, "one element column per thread");
  constexpr int RS = kPlaneRow<T, TK>;
  if (t < 0 || t >= sg.n_tiles) return;  // block-uniform: the whole block skips every barrier
  const int njk = sg.nj * sg.nk;
  const int jk = t % njk, is = t / njk;
  const int i0 = is * sg.L;
  const int j0 = (jk / sg.nk) * (TJ - 1);
  const int k0 = (jk % sg.nk) * (TK - 1);
  const int tid = threadIdx.x, tj = tid / TK, tk = tid % TK;
  const int j = j0 - 1 + tj, k = k0 - 1 + tk;
  // the extra node of the plane: row TJ (TK + 1 nodes) and column TK (TJ nodes)
  const bool has_x = tid < TJ + TK + 1;
  const int xj = tid < TK + 1 ? TJ : tid - (TK + 1), xk = tid < TK + 1 ? tid : TK;
  const int jx = j0 - 1 + xj, kx = k0 - 1 + xk;
  const bool col_in = j >= 0 && j <= g.ny && k >= 0 && k <= g.nz;
  const bool x_in = has_x && jx >= 0 && jx <= g.ny && kx >= 0 && kx <= g.nz;
  const bool own_col = tj >= 1 && tk >= 1 && j <= g.ny && k <= g.nz;
  const bool el_col = j >= 0 && j < g.ny && k >= 0 && k < g.nz;
  const int elast = min(i0 + sg.L - 1, g.nx);  // last owned node plane

  using Raw = decltype(ld((int64_t)0));
  Raw ra[3], rx[3];
  bool oka = false, okx = false;
  auto issue = [&](int pl) {
    oka = col_in && pl >= 0 && pl <= g.nx;
    okx = x_in && pl >= 0 && pl <= g.nx;
    const int64_t qa = 3 * (((int64_t)pl * g.nny + j) * g.nnz + k);
    const int64_t qx = 3 * (((int64_t)pl * g.nny + jx) * g.nnz + kx);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (oka) ra[c] = ld(qa + c);
      if (okx) rx[c] = ld(qx + c);
    }
  };
  auto store_plane = [&](int par) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      sm.plane[par][c][tj * RS + tk] = oka ? val(ra[c]) : T(0);
      if (has_x) sm.plane[par][c][xj * RS + xk] = okx ? val(rx[c]) : T(0);
    }
  };
  auto scale = [&](int e) {
    return (el_col && e >= 0 && e < g.nx) ? s[((int64_t)e * g.ny + j) * g.nz + k] : T(0);
  };

  // node (e, j, k) is corner (oy, oz) of the element columns (j - oy, k - oz): its four partial
  // sums from slot buffer sb, in a fixed order
  using Pre = decltype(pre((int64_t)-1));
  auto finish = [&](int sb, int64_t n, const T (&own)[3], const T (&xo)[3], const Pre& pv) {
    T y[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      y[c] = own[c];
      y[c] += sm.slot[sb][0 * 3 + c][tid - TK];
      y[c] += sm.slot[sb][1 * 3 + c][tid - 1];
      y[c] += sm.slot[sb][2 * 3 + c][tid - TK - 1];
    }
    fin(n, y, xo, pv);
  };

  issue(i0 - 1);
  store_plane(0);
  issue(i0);
  T sv = scale(i0 - 1);
  T prev1[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int c = 0; c < 3; ++c) prev1[a][c] = T(0);
  // one-barrier mode: the node of the previous step, finished after this step's barrier
  Pre pv_h{};
  T xo_h[3] = {T(0), T(0), T(0)}, own_h[3] = {T(0), T(0), T(0)};
  int64_t n_h = 0;
  bool out_h = false;

  int pb = 0;  // buffer holding node plane e
  for (int e = i0 - 1, st = 0; e <= elast; ++e, ++st) {
    const int pn = kOneBarrier<T> ? (pb == 2 ? 0 : pb + 1) : (pb ^ 1);
    store_plane(pn);  // node plane e + 1
    // Vector memory operations complete in issue order for the wait counter: the epilogue's
    // loads (waited for when the node is finished) and the next scale go out before the plane
    // two steps ahead, so waiting for them does not wait for that prefetch.
    const bool outp = e >= i0 && own_col;
    const int64_t n = ((int64_t)e * g.nny + j) * g.nnz + k;
    auto pv = pre(outp ? n : (int64_t)-1);
    const T se = sv;
    if (e + 1 <= elast) sv = scale(e + 1);
    if (e + 2 <= elast + 1) issue(e + 2);
    // Two-barrier mode: plane e + 1 is stored. One-barrier mode also: the previous step's slot
    // buffer is written, and every thread is past its reads of the plane buffer that the next
    // step overwrites (three buffers) and of the slot buffer that this step overwrites (two).
    __syncthreads();
    if (kOneBarrier<T> && out_h) finish((st - 1) & 1, n_h, own_h, xo_h, pv_h);
    T u[3][8], xo[3];
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int bx = b & 1, by = (b >> 1) & 1, bz = b >> 2;
      const int o = (tj + by) * RS + tk + bz;
#pragma unroll
      for (int c = 0; c < 3; ++c) u[c][b] = sm.plane[bx ? pn : pb][c][o];
    }
    // corner 0 is the thread's own node of plane e; taken here, before the barrier after which
    // a faster thread may overwrite this plane buffer
#pragma unroll
    for (int c = 0; c < 3; ++c) xo[c] = u[c][0];
    had_apply(u, K, se);
    const int sb = kOneBarrier<T> ? (st & 1) : 0;
    T own[3];
#pragma unroll
    for (int a = 0; a < 4; ++a)  // a = oy + 2 oz: corners b = 2a (ox = 0) and 2a + 1 (ox = 1)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T v = u[c][2 * a] + prev1[a][c];
        if (a == 0)
          own[c] = v;
        else
          sm.slot[sb][(a - 1) * 3 + c][tid] = v;
        prev1[a][c] = u[c][2 * a + 1];
      }
    if constexpr (kOneBarrier<T>) {
      out_h = outp;
      n_h = n;
      pv_h = pv;
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        xo_h[c] = xo[c];
        own_h[c] = own[c];
      }
    } else {
      __syncthreads();
      if (outp) finish(0, n, own, xo, pv);
    }
    pb = pn;
  }
  if constexpr (kOneBarrier<T>) {
    __syncthreads();
    if (out_h) finish((elast - i0 + 1) & 1, n_h, own_h, xo_h, pv_h);
  }
}

// The tile of this block: block b runs on XCD b % 8 and takes tile (b % 8) * per_xcd + b / 8.
// has = false runs no planes.
template <typename T, int TJ, int TK, typename LD, typename VAL, typename PRE, typename FIN>
__device__ __forceinline__ void stream_tile(StreamSmem<T, TJ, TK>& sm, const GridDims& g,
                                            const StreamGrid& sg, const HadK<T>& K,
                                            const T* __restrict__ s, bool has, LD ld, VAL val,
                                            PRE pre, FIN fin) {
  const int b = blockIdx.x;
  stream_tile_at(sm, g, sg, K, s, has ? (b & 7) * sg.per_xcd + (b >> 3) : -1, ld, val, pre,
                 fin);
}

// Launch shapes the streaming kernels are instantiated for (threads along j x along k).
#define AMT_STREAM_SHAPES(X) X(16, 16) X(8, 32) X(4, 64)

}  // namespace amt
