// tile_op.hpp — LDS-tiled matrix-free hex8 operator for the fine voxel grid.
//
// A block of 256 threads owns a TI x TJ x TK = 4 x 4 x 16 tile of nodes (one node per thread,
// k fastest as in the DOF numbering). It loads the tile plus a one-node halo (6 x 6 x 18
// nodes, 3 components each) into LDS with contiguous, coalesced loads along k, applying the
// caller's transform (e.g. p = z + beta*p_old) exactly once per node, and the 5 x 5 x 17
// element scales around the tile. Every thread then reads its 27 neighbours from LDS at
// compile-time offsets and applies Ke0 (uniform scalar loads). Compared with each thread
// gathering its 27 neighbours from global memory, this removes the per-neighbour 64-bit
// address arithmetic and bounds checks and the up-to-27-fold recomputation of transformed
// values.
//
// Block -> tile mapping is XCD-aware: the hardware deals consecutive workgroups round-robin
// to the 8 XCDs, so workgroup b is given tile (b % 8) * per_xcd + b / 8, which hands every
// XCD a contiguous slab of tiles whose halos then hit in that XCD's L2.
// Included from kernels.hpp after the operator primitives (GridDims, Nbr, apply_node).
#pragma once

namespace amt {

// LDS banking (32 banks of 4 bytes; a wave's 64-bit reads are served 16 lanes per clock, its
// 32-bit reads 32 lanes per clock). The 16 lanes of one k-row read 16 consecutive nodes, which
// at 3 values per node hit 16 distinct bank pairs in FP64, so 64-bit reads are conflict-free.
// In FP32 a 32-lane group holds two k-rows: tile_thread puts rows j and j + 2 in one group and
// the FP32 layout pads the LDS row to 24 nodes, so the two rows sit 3 * 48 = 144 = 16 (mod 32)
// banks apart and the group's 32 reads hit the 32 banks once. (The 4 x 8 x 8 tile put two
// rows of 8 in each 16-lane group and took two clocks per FP64 read.)
constexpr int kTI = 4, kTJ = 4, kTK = 16;
constexpr int kHI = kTI + 2, kHJ = kTJ + 2, kHK = kTK + 2;
constexpr int kHalo = kHI * kHJ * kHK;               // 648 nodes
constexpr int kEI = kTI + 1, kEJ = kTJ + 1, kEK = kTK + 1;
constexpr int kETile = kEI * kEJ * kEK;              // 425 elements
static_assert(kTI * kTJ * kTK == kBlock, "one node per thread");
static_assert(kTK == 16 && kTJ == 4, "tile_thread's row pairing and the FP32 padding assume it");

// LDS row strides (in nodes / elements) and sizes of the tile arrays
template <typename T> constexpr int kRowN = sizeof(T) == 4 ? 24 : kHK;
template <typename T> constexpr int kRowE = sizeof(T) == 4 ? 24 : kEK;
template <typename T> constexpr int kHaloL = kHI * kHJ * kRowN<T>;
template <typename T> constexpr int kETileL = kEI * kEJ * kRowE<T>;

struct TileGrid {
  int ti, tj, tk;   // tiles per axis
  int n_tiles;
  int per_xcd;      // ceil(n_tiles / 8)
};

inline TileGrid make_tile_grid(const GridDims& g) {
  TileGrid t;
  t.ti = (g.nx + 1 + kTI - 1) / kTI;
  t.tj = (g.ny + 1 + kTJ - 1) / kTJ;
  t.tk = (g.nz + 1 + kTK - 1) / kTK;
  t.n_tiles = t.ti * t.tj * t.tk;
  t.per_xcd = (t.n_tiles + 7) / 8;
  return t;
}
// launch with 8 * per_xcd blocks; blocks mapped past n_tiles exit after the barriers

template <typename T>
struct TileSmem {
  T x[kHaloL<T> * 3];  // node (hi, hj, hk) at (hi * kHJ + hj) * kRowN + hk
  T s[kETileL<T>];     // element (ei, ej, ek) at (ei * kEJ + ej) * kRowE + ek
};

// Tile origin (node coordinates) of this block; returns false if the block has no tile.
__device__ __forceinline__ bool tile_origin(const TileGrid& tg, int& i0, int& j0, int& k0) {
  const int b = blockIdx.x;
  const int t = (b & 7) * tg.per_xcd + (b >> 3);
  if (t >= tg.n_tiles) return false;
  const int tk = t % tg.tk, tj = (t / tg.tk) % tg.tj, ti = t / (tg.tk * tg.tj);
  i0 = ti * kTI;
  j0 = tj * kTJ;
  k0 = tk * kTK;
  return true;
}

// Cooperative load: halo node values f(dof) (0 outside the grid) and element scales.
// Fully unrolled, with every global load of the thread issued before the first LDS store, so
// the block pays the memory latency once instead of once per loop trip (the loads are
// independent; a rolled loop serialises them behind its stores).
constexpr int kNPer = (kHalo + kBlock - 1) / kBlock;   // halo nodes per thread (3)
constexpr int kSPer = (kETile + kBlock - 1) / kBlock;  // element scales per thread (2)

// Thread slots are whole halo nodes (three values each): the slot's tile coordinates are
// decoded once per node, not once per value, and its global offset is a 32-bit offset from
// the tile's (uniform, 64-bit) base, so the per-slot address arithmetic stays small.
// Halo slot t of this thread: its first DOF q and whether the node lies in the grid.
__device__ __forceinline__ bool tile_slot(const GridDims& g, int i0, int j0, int k0, int t,
                                          int64_t& q) {
  const int64_t nbase = ((int64_t)(i0 - 1) * g.nny + (j0 - 1)) * g.nnz + (k0 - 1);
  const int node = threadIdx.x + t * kBlock;
  const int hk = node % kHK, hj = (node / kHK) % kHJ, hi = node / (kHK * kHJ);
  const int i = i0 + hi - 1, j = j0 + hj - 1, k = k0 + hk - 1;
  q = 3 * (nbase + (hi * g.nny + hj) * g.nnz + hk);
  return node < kHalo && i >= 0 && i <= g.nx && j >= 0 && j <= g.ny && k >= 0 && k <= g.nz;
}

// Element-scale slot t of this thread (0 outside the grid).
template <typename T>
__device__ __forceinline__ T tile_scale(const GridDims& g, int i0, int j0, int k0, int t,
                                        const T* __restrict__ s) {
  const int64_t ebase = ((int64_t)(i0 - 1) * g.ny + (j0 - 1)) * g.nz + (k0 - 1);
  const int q = threadIdx.x + t * kBlock;
  const int ek = q % kEK, ej = (q / kEK) % kEJ, ei = q / (kEK * kEJ);
  const int i = i0 + ei - 1, j = j0 + ej - 1, k = k0 + ek - 1;
  const bool ok = q < kETile && i >= 0 && i < g.nx && j >= 0 && j < g.ny && k >= 0 && k < g.nz;
  return ok ? s[ebase + (ei * g.ny + ej) * g.nz + ek] : T(0);
}

template <typename T, typename F>
__device__ __forceinline__ void tile_gather(T (&v)[kNPer][3], T (&sv)[kSPer], const GridDims& g,
                                            int i0, int j0, int k0, F f,
                                            const T* __restrict__ s) {
#pragma unroll
  for (int t = 0; t < kNPer; ++t) {
    int64_t q;
    const bool ok = tile_slot(g, i0, j0, k0, t, q);
#pragma unroll
    for (int c = 0; c < 3; ++c) v[t][c] = ok ? f(q + c) : T(0);
  }
#pragma unroll
  for (int t = 0; t < kSPer; ++t) sv[t] = tile_scale(g, i0, j0, k0, t, s);
}

// LDS stores of gathered halo values and element scales (the layout of TileSmem).
template <typename T>
__device__ __forceinline__ void tile_store(TileSmem<T>& sm, const T (&v)[kNPer][3],
                                           const T (&sv)[kSPer]) {
#pragma unroll
  for (int t = 0; t < kNPer; ++t) {
    const int node = threadIdx.x + t * kBlock;
    if (node < kHalo) {
      const int l = (node / kHK) * kRowN<T> + node % kHK;
#pragma unroll
      for (int c = 0; c < 3; ++c) sm.x[3 * l + c] = v[t][c];
    }
  }
#pragma unroll
  for (int t = 0; t < kSPer; ++t) {
    const int q = threadIdx.x + t * kBlock;
    if (q < kETile) sm.s[(q / kEK) * kRowE<T> + q % kEK] = sv[t];
  }
}

template <typename T, typename F>
__device__ __forceinline__ void tile_load(TileSmem<T>& sm, const GridDims& g, int i0, int j0,
                                          int k0, F f, const T* __restrict__ s) {
  T v[kNPer][3], sv[kSPer];
  tile_gather(v, sv, g, i0, j0, k0, f, s);
  tile_store(sm, v, sv);
}

// This thread's node inside the tile: a wave is one i-plane, its four 16-lane rows hold
// j = 0, 2, 1, 3 (rows j and j + 2 share a 32-lane LDS group).
__device__ __forceinline__ void tile_thread(int& li, int& lj, int& lk) {
  lk = threadIdx.x % kTK;
  const int r = (threadIdx.x / kTK) % kTJ;
  lj = ((r & 1) << 1) | (r >> 1);
  li = threadIdx.x / (kTK * kTJ);
}

// Neighbourhood and element scales of the thread's node from LDS.
template <typename T>
__device__ __forceinline__ void tile_fetch(const TileSmem<T>& sm, int li, int lj, int lk,
                                           Nbr<T>& nb, T (&s8)[8]) {
  const int base = ((li + 1) * kHJ + (lj + 1)) * kRowN<T> + (lk + 1);
#pragma unroll
  for (int di = -1; di <= 1; ++di)
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
      for (int dk = -1; dk <= 1; ++dk) {
        const int m = (di + 1) * 9 + (dj + 1) * 3 + (dk + 1);
        const int o = 3 * (base + (di * kHJ + dj) * kRowN<T> + dk);
#pragma unroll
        for (int c = 0; c < 3; ++c) nb.v[m][c] = sm.x[o + c];
      }
  // element (i - ox[a], ...) sits at tile element index (li+1-ox, lj+1-oy, lk+1-oz)
  const int eb = ((li + 1) * kEJ + (lj + 1)) * kRowE<T> + (lk + 1);
#pragma unroll
  for (int a = 0; a < 8; ++a) s8[a] = sm.s[eb - (kOX[a] * kEJ + kOY[a]) * kRowE<T> - kOZ[a]];
}

// Element-loop form of tile_fetch + apply_node: y = sum_a s_a Ke0[3a.., :] x_e(a), reading
// each element's 24 nodal values from LDS as it is used (the empty asm with a memory clobber
// keeps the compiler from merging the eight elements' loads into one 81-value live set), so
// a thread holds 24 values instead of 81: more waves per SIMD for more LDS reads. own: the
// thread's node value.
template <typename T, typename TK>
__device__ __forceinline__ void tile_apply(const TileSmem<T>& sm, int li, int lj, int lk,
                                           const TK* __restrict__ Ke, T (&y)[3], T (&own)[3]) {
  const int nb0 = ((li + 1) * kHJ + (lj + 1)) * kRowN<T> + (lk + 1);
  const int eb = ((li + 1) * kEJ + (lj + 1)) * kRowE<T> + (lk + 1);
  y[0] = y[1] = y[2] = T(0);
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    asm volatile("" ::: "memory");
    const T se = sm.s[eb - (kOX[a] * kEJ + kOY[a]) * kRowE<T> - kOZ[a]];
    T t0 = T(0), t1 = T(0), t2 = T(0);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const int dn =
          ((kOX[b] - kOX[a]) * kHJ + (kOY[b] - kOY[a])) * kRowN<T> + (kOZ[b] - kOZ[a]);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        const T xv = sm.x[3 * (nb0 + dn) + d];
        t0 = fma((T)Ke[(3 * a + 0) * 24 + 3 * b + d], xv, t0);
        t1 = fma((T)Ke[(3 * a + 1) * 24 + 3 * b + d], xv, t1);
        t2 = fma((T)Ke[(3 * a + 2) * 24 + 3 * b + d], xv, t2);
      }
    }
    y[0] = fma(se, t0, y[0]);
    y[1] = fma(se, t1, y[1]);
    y[2] = fma(se, t2, y[2]);
  }
  asm volatile("" ::: "memory");
#pragma unroll
  for (int c = 0; c < 3; ++c) own[c] = sm.x[3 * nb0 + c];
}

}  // namespace amt
