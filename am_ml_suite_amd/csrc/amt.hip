// amt.hip — host orchestration and the C-ABI of libamt.so (declared in include/amt.h).
//
// One amt_problem owns every device buffer of one SIMP problem and one HIP stream. An OC
// iteration (topology_optimizer.cpp:294-384) runs entirely on the device:
//   k_simp -> k_diag -> k_dinv -> k_bc_rhs            ("assembly": operator coefficients)
//   PCG: k_pcg_init, then k_pcg_apply/k_pcg_update     ("solve"; state stays on the device,
//        pairs launched in chunks, host polls one flag  the host polls a done flag per chunk)
//   k_dot (compliance), k_sens                         ("sensitivity")
//   k_filter, k_oc_sweep x ~16, k_oc_final             ("update")
// and one small device->host copy of scalars ends it.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/amt.h"
#include "host_util.hpp"
#include "kernels.hpp"
#include "multigrid.hpp"
#include "mfma_op.hpp"
#include "stream_kernels.hpp"
#include "unet_api.hpp"
#include "unet_glue.hpp"
#include <cstddef>

namespace amt {
void hex8_stiffness(double E, double nu, double h, double* Ke);
}

namespace {
thread_local std::string g_create_error;
}

int amt::set_global_error(int code, const std::string& msg) {
  g_create_error = msg;
  return code;
}

using namespace amt;

namespace {

constexpr int kOcLevels = 3;  // bisection levels evaluated per OC sweep launch
constexpr int kOcBlocks = 512;  // OC sweep grid (partials per candidate)
constexpr int kOcCand = (1 << kOcLevels) - 1;
constexpr int kMaxReduceBlocks = 1024;
constexpr int kChunk = 32;  // PCG iterations launched between two host polls
// Largest tile grid for which the flexible-CG update runs inside the level-0 pre-smoothing
// (k_mg0_pre_upd): 2048 blocks each reading 2048 partials is 33 MB of L2 reads per launch; at
// 256x128x128 (19,305 tiles) it would be 3 GB, so larger grids keep k_fcg_update.
constexpr unsigned kFusedUpdateTiles = 2048;
constexpr int kMgChunk = 2;  // flexible-CG iterations per captured graph (even)
constexpr int kNnChunk = 4;      // learned FCG: iterations between host polls (even)
constexpr int kNnPatience = 16;  // learned FCG: chunks allowed without a 10% |r| drop
constexpr double kNnGainMomentum = 0.05;  // running means behind the correction's weight
constexpr float kMgOmega = 0.6f;  // damped-Jacobi weight of the V-cycle smoother...
constexpr float kMgTheta = 1.9f;  // ...capped per row at theta / sum_j |A_ij| (spectrum < 2)
// Chebyshev smoother of the levels >= 2: polynomial in M^-1 A with M = diag(row sums of |A|),
// whose spectrum lies in (0, 1] (Gershgorin); the polynomial damps [1 / kChebRatio, 1].
// Step s (from 1): d' = a_s d + b_s M^-1 (rhs - A x), x' = x + d' (the three-term recurrence
// of the Chebyshev iteration, with a_1 = 0).
constexpr double kChebRatio = 160.0;
// On grids of more than kChebExtraMinElements elements the two stencil levels above the dense
// one take kChebCoarseExtra Chebyshev steps more than mg_coarse_cheb: on the 120x60x60 bracket
// degree 4 there instead of 3 cut the FCG iterations per OC iteration from 33.1 to 29.6 for
// two more ~4 us launches per V-cycle and level. Problems of up to 64^3 elements converge in
// ~17 iterations either way, and the extra launches only cost them (64-problem set, 16^3-64^3:
// 15.0 against 13.9 OC iter/s).
constexpr int kChebCoarseExtra = 1;
// Damped-Jacobi stencil levels of more than this many nodes sweep half-precision blocks
// (k_mg_stencil_f16). A sweep streams the level's stored stencil (504 B per node in FP32): at
// 256x128x128 level 1 (545k nodes, 275 MB) that comes from HBM, and halving it gave 19.6 ->
// 20.1 OC iter/s; at 120x60x60 (59k nodes, 30 MB) the stencil stays in the caches, the sweep is
// bound by its load instructions, and the scale loads made it slower (137.0 -> 135.1).
constexpr int64_t kHalfStencilMinNodes = 1 << 18;
constexpr int64_t kChebExtraMinElements = 1 << 18;
struct ChebCoef {
  double a, b;
};
inline ChebCoef cheb_coef(int step) {
  const double lmax = 1.0, lmin = 1.0 / kChebRatio;
  const double th = 0.5 * (lmax + lmin), de = 0.5 * (lmax - lmin), sigma = th / de;
  double rho = 1.0 / sigma;
  if (step <= 1) return {0.0, 1.0 / th};
  ChebCoef c{0.0, 0.0};
  for (int k = 2; k <= step; ++k) {
    const double rn = 1.0 / (2.0 * sigma - rho);
    c = {rn * rho, 2.0 * rn / de};
    rho = rn;
  }
  return c;
}
// coarsen until a level has at most this many DOFs; the coarsest level's dense inverse runs in
// one block with up to 6 x 6 register-resident entries per thread (n <= 192)
constexpr int64_t kMgCoarseDofs = 192;

// Level-0 operator candidates: 0 = the 4 x 4 x 16 node gather tile (tile_op.hpp); 1 + 3 s + l =
// the element-column streaming tile (stream_op.hpp) of shape kStreamShape[s] (threads along j,
// k) with kStreamPlanes[l] node planes per block (4, 8, 16); 10 + 2 s + l = the same shapes with
// kStreamPlanes[3 + l] planes (2, 3).
constexpr int kStreamShape[3][2] = {{16, 16}, {8, 32}, {4, 64}};
constexpr int kStreamPlanes[5] = {4, 8, 16, 2, 3};
constexpr int kOpCandidates = 1 + 3 * 5;
inline void stream_candidate(int cand, int& shape, int& planes) {
  if (cand <= 9) {
    shape = (cand - 1) / 3;
    planes = kStreamPlanes[(cand - 1) % 3];
  } else {
    shape = (cand - 10) / 2;
    planes = kStreamPlanes[3 + (cand - 10) % 2];
  }
}

// Partials regions (doubles) inside one buffer; see amt_problem::part()
enum Region { R_APPLY = 0, R_UPD, R_INIT, R_OC, R_MISC, R_COUNT };

// Device scalar slots
enum Scal { S_MAXDIAG = 0, S_P, S_COMPL, S_NDESIGN, S_RR, S_BB, S_FLAG, S_NNZZ, S_NNRDR,
            S_NNUAU, S_NNUR, S_NND0, S_NNTH = S_NND0 + 5, S_PROJ, S_COUNT = S_PROJ + 2 };


}  // namespace

struct amt_problem {
  amt_config cfg;
  GridDims g;
  TileGrid tg;
  unsigned nb_tile = 0;  // blocks of the LDS-tiled operator kernels
  // level-0 operator launch chosen per kernel (autotuned on the first multigrid solve)
  struct OpSel {
    int cand = 0;    // index into the candidates above
    int shape = -1;  // stream shape, -1 for the gather tile
    StreamGrid sg;
    unsigned nb = 0;
    double us = 0.0;  // measured launch time (0 if forced or untuned)
  };
  OpSel op_apply, op_upd, op_post;  // FP64 CG operator / FP32 pre-smoothing / FP32 post
  bool ops_tuned = false;
  double tune_us[3][kOpCandidates] = {{0.0}};  // each candidate's measured time (0: not run)
  HadK<double> had64;
  HadK<float> had32;
  FcgState* tune_st = nullptr;
  int64_t n_e = 0, n_n = 0, n_dof = 0;
  hipStream_t stream = nullptr;
  std::string err;

  double* Ke64 = nullptr;
  float* Ke32 = nullptr;
  uint16_t *Kh16 = nullptr, *Kl16 = nullptr;  // split-BF16 Ke0 of the MFMA operator
  double *rho = nullptr, *s64 = nullptr;
  float* s32 = nullptr;
  uint8_t *design = nullptr, *fixcnt = nullptr;
  double *F = nullptr, *Fb = nullptr, *diag = nullptr, *dinv = nullptr;
  float* dinv32 = nullptr;
  double *u = nullptr, *r = nullptr, *z = nullptr, *p[2] = {nullptr, nullptr}, *q = nullptr;
  double* u2 = nullptr;  // warm_start 2: the displacement before the previous one (zeros at first)
  float *r32 = nullptr, *z32 = nullptr, *p32[2] = {nullptr, nullptr}, *q32 = nullptr,
        *d32 = nullptr;
  double *dc = nullptr, *dcf = nullptr;
  StencilEntry* stencil = nullptr;
  double* stencil_w = nullptr;
  int n_stencil = 0;
  int stencil_reach = 0;  // max |offset| of the stencil (the LDS-tiled filter takes <= 3)
  double* partials = nullptr;
  int64_t region_off[R_COUNT] = {0};
  int stride_apply = 0, stride_red = 0;
  double* part(Region r) { return partials + region_off[r]; }
  double* scal = nullptr;
  PcgState* pcg = nullptr;   // [2], alternating by iteration parity
  PcgState* pcg_flag = nullptr;
  OcState* oc = nullptr;
  int64_t* dbg_idx = nullptr;
  double* dbg_w = nullptr;
  int* dbg_cnt = nullptr;

  // geometric multigrid (precond == AMT_PRECOND_MULTIGRID)
  struct MgLevel {
    MgGrid g;
    MgBrick brick{};     // 4x4x4-node bricks of the fused kernels (levels >= 1)
    float* A = nullptr;  // 27-point block stencil (levels >= 1)
    // half-precision blocks with per-node scales for the one-launch-per-sweep damped-Jacobi
    // levels of more than kHalfStencilMinNodes nodes (k_mg_stencil_f16; level 1 in practice)
    _Float16* Ah = nullptr;
    float* Ahs = nullptr;
    float *wd = nullptr, *rhs = nullptr, *x = nullptr, *t = nullptr;
    // Chebyshev levels: 1 / row sums and five iterate / direction buffers
    float* mi = nullptr;
    float* cx[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  };
  std::vector<MgLevel> mg;
  int mg_fuse_from = 0;       // first level of the fused V-cycle kernels (0: none)
  double* mg_G = nullptr;     // 8 x 576 Galerkin child tables
  float* mg_ainv = nullptr;   // coarsest dense inverse
  int mg_nc = 0;              // coarsest DOFs
  FcgState* fcg = nullptr;
  double* rr[2] = {nullptr, nullptr};  // flexible-CG residual, alternating (rr[0] is r)
  // run_iteration reads the solve's true residual together with the OC state and compliance in
  // one copy at the end (defer_true_rr set while it solves; pending_rhs2 = |b|^2 of that solve)
  bool defer_true_rr = false;
  double pending_rhs2 = 0.0;
  IterScalars* iter_scal = nullptr;
  // one captured chunk of kMgChunk flexible-CG iterations, replayed per chunk
  hipGraph_t fcg_graph = nullptr;
  // the post-solve launches of an OC iteration, captured once
  hipGraph_t post_graph = nullptr;
  hipGraphExec_t post_exec = nullptr;
  bool post_graph_failed = false;
  hipGraphExec_t fcg_exec = nullptr;
  bool fcg_graph_failed = false;
  int mg_last_iters = 0;  // iterations of the previous solve: sizes the first launch batch


  // learned preconditioner (precond == AMT_PRECOND_UNET, or a learned warm start)
  amt_unet* net = nullptr;  // attached, not owned
  float* nn_feat = nullptr;
  float *nn_U = nullptr, *nn_AU = nullptr, *nn_J = nullptr, *nn_KJ = nullptr;
  // training samples: a residual (FP32) and the density it belongs to (the sample's operator
  // and density feature are rebuilt from it when the network trains on it)
  struct Sample {
    float* r;
    double* rho;
  };
  std::vector<Sample> bank;
  int warm_collect = 0;  // multigrid warm solves record their start residual while bank < this

  bool loads_set = false, design_ready = false, coeffs_ready = false, have_u = false;
  double n_design = 0;

  // operator timing (HIP events around k_pcg_apply launches)
  bool prof_on = false;
  std::vector<hipEvent_t> ev_pool;
  int ev_used = 0;
  std::vector<int> ev_kind;  // per used event pair: 0 operator, 1 level-0 pre-smoothing + update
  int64_t prof_launches = 0, prof_launches_upd = 0;
  double prof_ms = 0.0, prof_ms_upd = 0.0;
  int prof_every = 8, prof_tick = 0;

  std::vector<void*> allocs;
};

namespace {

int fail(amt_problem* p, int code, const std::string& msg) {
  if (p)
    p->err = msg;
  else
    g_create_error = msg;
  return code;
}

#define HIP_TRY(p, call)                                                              \
  do {                                                                                \
    hipError_t e_ = (call);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail((p), AMT_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define AMT_CHECK(call)            \
  do {                             \
    int rc_ = (call);              \
    if (rc_ != AMT_OK) return rc_; \
  } while (0)

template <typename T>
int dalloc(amt_problem* p, T** ptr, size_t count) {
  void* v = nullptr;
  hipError_t e = hipMalloc(&v, std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess)
    return fail(p, AMT_EHIP, std::string("hipMalloc: ") + hipGetErrorString(e));
  p->allocs.push_back(v);
  *ptr = (T*)v;
  return AMT_OK;
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }
inline unsigned nred(int64_t n) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(nblk(n), kMaxReduceBlocks));
}

int check_launch(amt_problem* p, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(p, AMT_EHIP, std::string(what) + ": " + hipGetErrorString(e));
  return AMT_OK;
}


int copy_in(amt_problem* p, void* dst, const void* src, size_t bytes) {
  HIP_TRY(p, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, p->stream));
  HIP_TRY(p, hipStreamSynchronize(p->stream));
  return AMT_OK;
}
int copy_out(amt_problem* p, void* dst, const void* src, size_t bytes) {
  HIP_TRY(p, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(p, hipStreamSynchronize(p->stream));
  return AMT_OK;
}

int mg_setup(amt_problem* p);
void dot_dev(amt_problem* p, const double* a, const double* b, double* out);
int bank_push(amt_problem* p, const double* r);
int learned_warm_start(amt_problem* p);
int residual64(amt_problem* p, const double* x, double* rout, double* rr_host);
void prof_pick(amt_problem* p, hipEvent_t* e0, hipEvent_t* e1);
void prof_take(amt_problem* p, int kind, hipEvent_t* e0, hipEvent_t* e1);
void prof_harvest(amt_problem* p);

// ---- operator coefficients for the current density ("assembly") ----
// dcf = filtered dc: the LDS-tiled kernel for stencils reaching at most 3 elements, the
// global-gather one beyond (same result bit for bit).
void launch_filter(amt_problem* p) {
  const int R = p->stencil_reach;
  if (R <= 3) {
    int ty, tz;
    filter_tile_dims(R, ty, tz);
    const GridDims& g = p->g;
    const int64_t blocks = (int64_t)((g.nx + kFilterTX - 1) / kFilterTX) * ((g.ny + ty - 1) / ty) *
                           ((g.nz + tz - 1) / tz);
    hipLaunchKernelGGL(k_filter_tile, dim3((unsigned)blocks), dim3(kBlock), 0, p->stream, p->dc,
                       p->rho, p->stencil, p->stencil_w, p->n_stencil, g, R, ty, tz, p->dcf);
  } else {
    hipLaunchKernelGGL(k_filter, dim3(nblk(p->n_e)), dim3(kBlock), 0, p->stream, p->dc, p->rho,
                       p->stencil, p->stencil_w, p->n_stencil, p->g, p->dcf);
  }
}

int prepare_coeffs(amt_problem* p) {
  const GridDims g = p->g;
  hipLaunchKernelGGL(k_simp, dim3(nblk(p->n_e)), dim3(kBlock), 0, p->stream, p->rho, p->s64,
                     p->s32, g, p->cfg.density_index == AMT_INDEX_REF_COMPAT ? 1 : 0,
                     p->cfg.penalty, p->cfg.E_min, p->cfg.rho_min);
  hipLaunchKernelGGL(k_diag, dim3(nblk(p->n_n)), dim3(kBlock), 0, p->stream, p->diag, p->s64,
                     p->Ke64, g, p->part(R_APPLY), p->stride_apply);
  hipLaunchKernelGGL((k_finalize<1, true>), dim3(1), dim3(kBlock), 0, p->stream,
                     p->part(R_APPLY), (int)nblk(p->n_n), p->stride_apply,
                     p->scal + S_MAXDIAG);
  hipLaunchKernelGGL(k_dinv, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream, p->diag,
                     p->fixcnt, p->scal + S_MAXDIAG, p->cfg.penalty_factor, p->scal + S_P,
                     p->dinv, p->dinv32, p->n_dof);
  hipLaunchKernelGGL(k_bc_rhs, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream, p->F,
                     p->fixcnt, p->Fb, p->n_dof);
  AMT_CHECK(check_launch(p, "prepare_coeffs"));
  if (!p->mg.empty()) AMT_CHECK(mg_setup(p));
  if (p->nn_feat)
    hipLaunchKernelGGL(k_nn_feature, dim3(nblk(p->n_n)), dim3(kBlock), 0, p->stream, p->s64,
                       p->g, p->nn_feat);
  p->coeffs_ready = true;
  return AMT_OK;
}

// ---- geometric multigrid: hierarchy, per-iteration coarse operators, V-cycle ----

// G_c = Q_c^T Ke0 Q_c for the 8 children c of a coarse element; Q_c interpolates the coarse
// element's 8 nodal vectors trilinearly to child c's 8 nodes (local node order of
// element_kernel.cpp:196-199).
void mg_child_tables(const double* Ke, double* G) {
  static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                                {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  for (int c = 0; c < 8; ++c) {
    double w[8][8];  // w[b][a]: weight of coarse node a at child node b
    for (int b = 0; b < 8; ++b)
      for (int a = 0; a < 8; ++a) {
        double v = 1.0;
        for (int k = 0; k < 3; ++k) {
          const double t = (off[c][k] + off[b][k]) / 2.0;
          v *= off[a][k] ? t : 1.0 - t;
        }
        w[b][a] = v;
      }
    // G[3a+i][3a'+j] = sum_{b,b'} w[b][a] Ke[3b+i][3b'+j] w[b'][a']
    for (int a = 0; a < 8; ++a)
      for (int i = 0; i < 3; ++i)
        for (int a2 = 0; a2 < 8; ++a2)
          for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
            for (int b = 0; b < 8; ++b) {
              if (w[b][a] == 0.0) continue;
              for (int b2 = 0; b2 < 8; ++b2)
                acc += w[b][a] * Ke[(3 * b + i) * 24 + 3 * b2 + j] * w[b2][a2];
            }
            G[c * 576 + (3 * a + i) * 24 + 3 * a2 + j] = acc;
          }
  }
}

MgGrid mg_grid(int ex, int ey, int ez) {
  MgGrid g;
  g.ex = ex;
  g.ey = ey;
  g.ez = ez;
  g.nnx = ex + 1;
  g.nny = ey + 1;
  g.nnz = ez + 1;
  g.n_nodes = (int64_t)g.nnx * g.nny * g.nnz;
  return g;
}

int mg_build_hierarchy(amt_problem* p, const double* Ke) {
  amt_problem::MgLevel L0;
  L0.g = mg_grid(p->g.nx, p->g.ny, p->g.nz);
  AMT_CHECK(dalloc(p, &L0.wd, p->n_dof));
  AMT_CHECK(dalloc(p, &L0.x, p->n_dof));
  AMT_CHECK(dalloc(p, &L0.t, p->n_dof));
  p->mg.push_back(L0);
  for (;;) {
    const MgGrid& f = p->mg.back().g;
    if (3 * f.n_nodes <= kMgCoarseDofs && p->mg.size() > 1) break;
    if (f.ex == 1 && f.ey == 1 && f.ez == 1 && p->mg.size() > 1) break;
    amt_problem::MgLevel L;
    L.g = mg_grid((f.ex + 1) / 2, (f.ey + 1) / 2, (f.ez + 1) / 2);
    L.brick.nbi = (L.g.nnx + kFB - 1) / kFB;
    L.brick.nbj = (L.g.nny + kFB - 1) / kFB;
    L.brick.nbk = (L.g.nnz + kFB - 1) / kFB;
    const int64_t nd = 3 * L.g.n_nodes;
    AMT_CHECK(dalloc(p, &L.A, (size_t)9 * kStencilBlocks * L.g.n_nodes));
    AMT_CHECK(dalloc(p, &L.wd, nd));
    AMT_CHECK(dalloc(p, &L.rhs, nd));
    AMT_CHECK(dalloc(p, &L.x, nd));
    AMT_CHECK(dalloc(p, &L.t, nd));
    p->mg.push_back(L);
  }
  p->rr[0] = p->r;
  AMT_CHECK(dalloc(p, &p->rr[1], p->n_dof));
  p->mg_nc = (int)(3 * p->mg.back().g.n_nodes);
  if (p->mg_nc > kMgCoarseDofs) return fail(p, AMT_EINVAL, "multigrid coarsest level too large");
  // fused kernels on the stencil levels of at most kFuseMaxNodes nodes (the levels shrink, so
  // they are a tail of the hierarchy), when there is a stencil level above the dense one
  p->mg_fuse_from = 0;
  const int nl = (int)p->mg.size();
  // Chebyshev smoother on the stencil levels from 2 on
  if (p->cfg.mg_coarse_cheb > 0)
    for (int l = 2; l <= nl - 2; ++l) {
      const int64_t nd = 3 * p->mg[l].g.n_nodes;
      AMT_CHECK(dalloc(p, &p->mg[l].mi, nd));
      for (auto& b : p->mg[l].cx) AMT_CHECK(dalloc(p, &b, nd));
    }
  for (int l = 1; l <= nl - 2; ++l)
    if (!p->mg[l].mi && p->mg[l].g.n_nodes > kHalfStencilMinNodes) {
      AMT_CHECK(dalloc(p, &p->mg[l].Ah, (size_t)9 * kStencilBlocks * p->mg[l].g.n_nodes));
      AMT_CHECK(dalloc(p, &p->mg[l].Ahs, (size_t)p->mg[l].g.n_nodes));
    }
  if (p->cfg.mg_fused && nl >= 3 && p->mg_nc <= kDenseMax) {
    for (int l = 1; l <= nl - 2; ++l)
      if (p->mg[l].g.n_nodes <= kFuseMaxNodes) {
        p->mg_fuse_from = l;
        break;
      }
  }
  AMT_CHECK(dalloc(p, &p->mg_ainv, (size_t)p->mg_nc * p->mg_nc));
  std::vector<double> G(8 * 576);
  mg_child_tables(Ke, G.data());
  AMT_CHECK(dalloc(p, &p->mg_G, G.size()));
  AMT_CHECK(copy_in(p, p->mg_G, G.data(), G.size() * sizeof(double)));
  return AMT_OK;
}

// Coarse operators for the current element scales and penalty (after k_dinv).
int mg_setup(amt_problem* p) {
  auto& mg = p->mg;
  hipLaunchKernelGGL(k_mg_wd0, dim3(nblk(p->n_n), 3), dim3(kBlock), 0, p->stream, p->s64,
                     p->Ke64, p->g, p->fixcnt, p->scal + S_P, kMgOmega, kMgTheta, mg[0].wd);
  hipLaunchKernelGGL(k_mg_galerkin1, dim3(nblk(mg[1].g.n_nodes), kStencilBlocks, 3), dim3(kBlock), 0,
                     p->stream, p->s64, p->g, p->mg_G, p->fixcnt, p->scal + S_P, mg[1].g,
                     mg[1].A);
  for (size_t l = 1; l < mg.size(); ++l) {
    if (l >= 2)
      hipLaunchKernelGGL(k_mg_coarsen,
                         dim3((unsigned)((mg[l].g.n_nodes + kCoarsenNodes - 1) / kCoarsenNodes),
                              kStencilBlocks, 3),
                         dim3(9 * kCoarsenNodes), 0,
                         p->stream, mg[l - 1].A, mg[l - 1].g, mg[l].g, mg[l].A);
    hipLaunchKernelGGL(k_mg_stencil_wdinv, dim3(nblk(3 * mg[l].g.n_nodes)), dim3(kBlock), 0,
                       p->stream, mg[l].A, mg[l].g, kMgOmega, kMgTheta, mg[l].wd, mg[l].mi,
                       (float)cheb_coef(1).b);
    if (mg[l].Ah)
      hipLaunchKernelGGL(k_mg_stencil_f16, dim3(nblk(mg[l].g.n_nodes)), dim3(kBlock), 0,
                         p->stream, mg[l].A, mg[l].g, mg[l].Ah, mg[l].Ahs);
  }
  // coarsest level: dense inverse on the device (no host round trip, the OC iteration stays
  // asynchronous)
  const auto& C = mg.back();
  const int n = p->mg_nc;
#define AMT_DENSE_INV(R)                                                                        \
  case R:                                                                                       \
    hipLaunchKernelGGL(k_mg_dense_inverse<R>, dim3(1), dim3(1024), 0, p->stream, C.A, C.g, n,   \
                       p->mg_ainv);                                                             \
    break;
  switch ((n + 31) / 32) {
    AMT_DENSE_INV(1)
    AMT_DENSE_INV(2)
    AMT_DENSE_INV(3)
    AMT_DENSE_INV(4)
    AMT_DENSE_INV(5)
    AMT_DENSE_INV(6)
    default:
      return fail(p, AMT_EINVAL, "multigrid coarsest level too large");
  }
#undef AMT_DENSE_INV
  return check_launch(p, "mg_setup");
}

// ---- level-0 operator launches: gather tile or streaming tile, chosen per kernel ----

amt_problem::OpSel make_op_sel(const amt_problem* p, int cand) {
  amt_problem::OpSel s;
  s.cand = cand;
  if (cand <= 0) {
    s.cand = 0;
    s.shape = -1;
    s.nb = p->nb_tile;
    return s;
  }
  int planes = 0;
  stream_candidate(cand, s.shape, planes);
  s.sg = make_stream_grid(p->g, kStreamShape[s.shape][0], kStreamShape[s.shape][1], planes);
  s.nb = stream_blocks(s.sg);
  return s;
}

// KER<TJ, TK> for the selection's streaming shape, on the handle's stream.
#define AMT_STREAM_LAUNCH(SEL, KER, ...)                                                       \
  do {                                                                                         \
    switch ((SEL).shape) {                                                                     \
      case 0:                                                                                  \
        hipLaunchKernelGGL((KER<16, 16>), dim3((SEL).nb), dim3(kBlock), 0, p->stream,          \
                           __VA_ARGS__);                                                       \
        break;                                                                                 \
      case 1:                                                                                  \
        hipLaunchKernelGGL((KER<8, 32>), dim3((SEL).nb), dim3(kBlock), 0, p->stream,           \
                           __VA_ARGS__);                                                       \
        break;                                                                                 \
      default:                                                                                 \
        hipLaunchKernelGGL((KER<4, 64>), dim3((SEL).nb), dim3(kBlock), 0, p->stream,           \
                           __VA_ARGS__);                                                       \
        break;                                                                                 \
    }                                                                                          \
  } while (0)

// AMT_STREAM_LAUNCH with the kernel's start and end timestamped into events E0, E1 by the
// GPU (hipExtLaunchKernelGGL), for the sampled operator launches.
#define AMT_STREAM_LAUNCH_EV(SEL, KER, E0, E1, ...)                                             \
  do {                                                                                         \
    switch ((SEL).shape) {                                                                     \
      case 0:                                                                                  \
        hipExtLaunchKernelGGL((KER<16, 16>), dim3((SEL).nb), dim3(kBlock), 0, p->stream, E0,   \
                              E1, 0, __VA_ARGS__);                                             \
        break;                                                                                 \
      case 1:                                                                                  \
        hipExtLaunchKernelGGL((KER<8, 32>), dim3((SEL).nb), dim3(kBlock), 0, p->stream, E0,    \
                              E1, 0, __VA_ARGS__);                                             \
        break;                                                                                 \
      default:                                                                                 \
        hipExtLaunchKernelGGL((KER<4, 64>), dim3((SEL).nb), dim3(kBlock), 0, p->stream, E0,    \
                              E1, 0, __VA_ARGS__);                                             \
        break;                                                                                 \
    }                                                                                          \
  } while (0)

// y = K x (+ penalty with P) in FP64 / FP32 with the selection's launch.
void launch_apply64(amt_problem* p, const amt_problem::OpSel& s, const double* x, double* y,
                    const double* P) {
  if (s.shape < 0)
    hipLaunchKernelGGL((k_apply<double, double>), dim3(s.nb), dim3(kBlock), 0, p->stream, x, y,
                       p->s64, p->Ke64, p->fixcnt, P, p->g, p->tg);
  else if (s.shape == 0)
    hipLaunchKernelGGL((k_apply_s<double, 16, 16>), dim3(s.nb), dim3(kBlock), 0, p->stream, x,
                       y, p->s64, p->had64, p->fixcnt, P, p->g, s.sg);
  else if (s.shape == 1)
    hipLaunchKernelGGL((k_apply_s<double, 8, 32>), dim3(s.nb), dim3(kBlock), 0, p->stream, x,
                       y, p->s64, p->had64, p->fixcnt, P, p->g, s.sg);
  else
    hipLaunchKernelGGL((k_apply_s<double, 4, 64>), dim3(s.nb), dim3(kBlock), 0, p->stream, x,
                       y, p->s64, p->had64, p->fixcnt, P, p->g, s.sg);
}
void launch_apply32(amt_problem* p, const amt_problem::OpSel& s, const float* x, float* y,
                    const double* P) {
  if (s.shape < 0)
    hipLaunchKernelGGL((k_apply<float, float>), dim3(s.nb), dim3(kBlock), 0, p->stream, x, y,
                       p->s32, p->Ke32, p->fixcnt, P, p->g, p->tg);
  else if (s.shape == 0)
    hipLaunchKernelGGL((k_apply_s<float, 16, 16>), dim3(s.nb), dim3(kBlock), 0, p->stream, x, y,
                       p->s32, p->had32, p->fixcnt, P, p->g, s.sg);
  else if (s.shape == 1)
    hipLaunchKernelGGL((k_apply_s<float, 8, 32>), dim3(s.nb), dim3(kBlock), 0, p->stream, x, y,
                       p->s32, p->had32, p->fixcnt, P, p->g, s.sg);
  else
    hipLaunchKernelGGL((k_apply_s<float, 4, 64>), dim3(s.nb), dim3(kBlock), 0, p->stream, x, y,
                       p->s32, p->had32, p->fixcnt, P, p->g, s.sg);
}

// r = Fb - K x with block partials of |r|^2 into R_APPLY; returns the block count.
unsigned launch_residual(amt_problem* p, const double* x, double* rout) {
  const auto& s = p->op_apply;
  if (s.shape < 0)
    hipLaunchKernelGGL(k_residual, dim3(s.nb), dim3(kBlock), 0, p->stream, x, p->Fb, rout,
                       p->s64, p->Ke64, p->fixcnt, p->scal + S_P, p->g, p->tg, p->part(R_APPLY));
  else
    AMT_STREAM_LAUNCH(s, k_residual_s, x, p->Fb, rout, p->s64, p->had64, p->fixcnt,
                      p->scal + S_P, p->g, s.sg, p->part(R_APPLY));
  return s.nb;
}

// p_new = z + beta p_old, q = K p_new, partials p.q into R_APPLY (flexible CG). With events
// e0, e1 the GPU timestamps the kernel's start and end into them.
void launch_fcg_apply(amt_problem* p, const amt_problem::OpSel& s, const double* z,
                      const double* pold, double* pnew, const FcgState* st,
                      hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  if (e0) {
    if (s.shape < 0)
      hipExtLaunchKernelGGL(k_fcg_apply, dim3(s.nb), dim3(kBlock), 0, p->stream, e0, e1, 0, z,
                            pold, pnew, p->q, p->s64, p->Ke64, p->fixcnt, p->scal + S_P, p->g,
                            p->tg, st, p->part(R_APPLY));
    else
      AMT_STREAM_LAUNCH_EV(s, k_fcg_apply_s, e0, e1, z, pold, pnew, p->q, p->s64, p->had64,
                           p->fixcnt, p->scal + S_P, p->g, s.sg, st, p->part(R_APPLY));
    return;
  }
  if (s.shape < 0)
    hipLaunchKernelGGL(k_fcg_apply, dim3(s.nb), dim3(kBlock), 0, p->stream, z, pold, pnew, p->q,
                       p->s64, p->Ke64, p->fixcnt, p->scal + S_P, p->g, p->tg, st,
                       p->part(R_APPLY));
  else
    AMT_STREAM_LAUNCH(s, k_fcg_apply_s, z, pold, pnew, p->q, p->s64, p->had64, p->fixcnt,
                      p->scal + S_P, p->g, s.sg, st, p->part(R_APPLY));
}

void launch_mg0_pre_upd(amt_problem* p, const amt_problem::OpSel& s, const double* r_prev,
                        double* rnew, const double* q, double* x, const double* pnew,
                        const FcgState* st, int nba, hipEvent_t e0 = nullptr,
                        hipEvent_t e1 = nullptr) {
  auto& L0 = p->mg[0];
  if (e0) {  // timestamped by the GPU at the kernel's start and end
    if (s.shape < 0)
      hipExtLaunchKernelGGL(k_mg0_pre_upd, dim3(s.nb), dim3(kBlock), 0, p->stream, e0, e1, 0,
                            r_prev, rnew, q, x, pnew, p->g, p->tg, p->s32, p->Ke32, p->fixcnt,
                            p->scal + S_P, L0.wd, L0.x, L0.t, st,
                            (const double*)p->part(R_APPLY), nba, p->part(R_UPD));
    else
      AMT_STREAM_LAUNCH_EV(s, k_mg0_pre_upd_s, e0, e1, r_prev, rnew, q, x, pnew, p->g, s.sg,
                           p->s32, p->had32, p->fixcnt, p->scal + S_P, L0.wd, L0.x, L0.t, st,
                           (const double*)p->part(R_APPLY), nba, p->part(R_UPD));
    return;
  }
  if (s.shape < 0)
    hipLaunchKernelGGL(k_mg0_pre_upd, dim3(s.nb), dim3(kBlock), 0, p->stream, r_prev, rnew, q, x,
                       pnew, p->g, p->tg, p->s32, p->Ke32, p->fixcnt, p->scal + S_P, L0.wd,
                       L0.x, L0.t, st, (const double*)p->part(R_APPLY), nba, p->part(R_UPD));
  else
    AMT_STREAM_LAUNCH(s, k_mg0_pre_upd_s, r_prev, rnew, q, x, pnew, p->g, s.sg, p->s32,
                      p->had32, p->fixcnt, p->scal + S_P, L0.wd, L0.x, L0.t, st,
                      (const double*)p->part(R_APPLY), nba, p->part(R_UPD));
}

void launch_mg0_pre(amt_problem* p, const amt_problem::OpSel& s, const double* r,
                    const int* done) {
  auto& L0 = p->mg[0];
  if (s.shape < 0)
    hipLaunchKernelGGL(k_mg0_pre, dim3(s.nb), dim3(kBlock), 0, p->stream, r, p->g, p->tg,
                       p->s32, p->Ke32, p->fixcnt, p->scal + S_P, L0.wd, L0.x, L0.t, done);
  else
    AMT_STREAM_LAUNCH(s, k_mg0_pre_s, r, p->g, s.sg, p->s32, p->had32, p->fixcnt,
                      p->scal + S_P, L0.wd, L0.x, L0.t, done);
}

void launch_mg0_post(amt_problem* p, const amt_problem::OpSel& s, const double* r, double* z,
                     const double* q, double* partials, int stride, const int* done) {
  auto& L0 = p->mg[0];
  if (s.shape < 0)
    hipLaunchKernelGGL(k_mg0_post, dim3(s.nb), dim3(kBlock), 0, p->stream, r, p->g, p->tg,
                       p->s32, p->Ke32, p->fixcnt, p->scal + S_P, L0.wd, L0.x, z, q, partials,
                       stride, done);
  else
    AMT_STREAM_LAUNCH(s, k_mg0_post_s, r, p->g, s.sg, p->s32, p->had32, p->fixcnt,
                      p->scal + S_P, L0.wd, L0.x, z, q, partials, stride, done);
}

// Blocks (partials) of the level-0 post-smoothing launch.
unsigned mg0_post_blocks(const amt_problem* p) {
  return p->cfg.precision == AMT_PREC_MIXED_BF16 ? p->nb_tile : p->op_post.nb;
}

// Tuned choices per grid (candidate and measured microseconds for the three kernels), shared by
// every handle of the process.
struct TuneEntry {
  int cand[3];
  double us[3];
  double all_us[3][kOpCandidates];
};
std::mutex g_tune_mutex;
std::map<std::tuple<int, int, int>, TuneEntry> g_tune_cache;

// Autotuner of the level-0 operator launches: every candidate of each of the three kernels of
// the multigrid flexible CG is launched 2 + 5 times on the handle's buffers before the solve
// starts (outputs go to vectors the solve overwrites; the displacement is not touched) and the
// fastest mean time wins. All streaming candidates compute bitwise identical operator values;
// the choice changes the speed, and the rounding of the block-wise inner products (and, against
// the gather tile, of the operator sum).
int tune_ops(amt_problem* p) {
  if (p->ops_tuned) return AMT_OK;
  const auto key = std::make_tuple(p->g.nx, p->g.ny, p->g.nz);
  // held through the tuning: handles of one grid (the batch engine's threads) share one choice
  std::lock_guard<std::mutex> lk(g_tune_mutex);
  {
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      amt_problem::OpSel* sel[3] = {&p->op_apply, &p->op_upd, &p->op_post};
      for (int k = 0; k < 3; ++k) {
        *sel[k] = make_op_sel(p, it->second.cand[k]);
        sel[k]->us = it->second.us[k];
      }
      std::memcpy(p->tune_us, it->second.all_us, sizeof(p->tune_us));
      p->ops_tuned = true;
      return AMT_OK;
    }
  }
  if (!p->tune_st) AMT_CHECK(dalloc(p, &p->tune_st, 1));
  FcgState ts;
  std::memset(&ts, 0, sizeof(ts));
  ts.rz = 1.0;
  ts.beta = 0.5;
  AMT_CHECK(copy_in(p, p->tune_st, &ts, sizeof(ts)));
  const int* tdone = (const int*)((const char*)p->tune_st + offsetof(FcgState, done));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_TRY(p, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(p, AMT_EHIP, "hipEventCreate failed");
  }
  TuneEntry te;
  int rc = AMT_OK;
  for (int kind = 0; kind < 3 && rc == AMT_OK; ++kind) {
    double best = 1e300;
    int best_c = 0;
    for (int c = 0; c < kOpCandidates && rc == AMT_OK; ++c) {
      const amt_problem::OpSel s = make_op_sel(p, c);
      auto launch = [&]() {
        if (kind == 0)
          launch_fcg_apply(p, s, p->z, p->p[0], p->p[1], p->tune_st);
        else if (kind == 1)
          launch_mg0_pre_upd(p, s, p->r, p->rr[1], p->q, p->z, p->p[1], p->tune_st,
                             (int)p->op_apply.nb);
        else
          launch_mg0_post(p, s, p->r, p->z, p->q, p->part(R_INIT), p->stride_red, tdone);
      };
      for (int w = 0; w < 2; ++w) launch();
      (void)hipEventRecord(e0, p->stream);
      for (int w = 0; w < 5; ++w) launch();
      (void)hipEventRecord(e1, p->stream);
      float ms = 0.f;
      if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
        rc = fail(p, AMT_EHIP, "operator autotuning launch failed");
        break;
      }
      const double us = 1e3 * ms / 5.0;
      te.all_us[kind][c] = us;
      if (us < best) {
        best = us;
        best_c = c;
      }
    }
    te.cand[kind] = best_c;
    te.us[kind] = best;
    if (kind == 0) {  // the later kinds reduce this launch's partials
      p->op_apply = make_op_sel(p, best_c);
      p->op_apply.us = best;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  AMT_CHECK(rc);
  p->op_upd = make_op_sel(p, te.cand[1]);
  p->op_upd.us = te.us[1];
  p->op_post = make_op_sel(p, te.cand[2]);
  p->op_post.us = te.us[2];
  std::memcpy(p->tune_us, te.all_us, sizeof(p->tune_us));
  p->ops_tuned = true;
  g_tune_cache[key] = te;
  return AMT_OK;
}

// z = V-cycle(r) (FP32 inside); writes partials of z.r and z.q (q may be null) to R_INIT.
// With r_prev, the flexible-CG update is folded into the level-0 pre-smoothing
// (k_mg0_pre_upd): r = r_prev - alpha q is formed there, x += alpha p_new, and the |r|^2
// partials go to R_UPD, one per tile block.
void mg_vcycle(amt_problem* p, const double* r, double* z, const double* q,
               const double* r_prev = nullptr, const double* p_new = nullptr,
               hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr) {
  auto& mg = p->mg;
  const int* done = (const int*)((const char*)p->fcg + offsetof(FcgState, done));
  const size_t L = mg.size();
  const bool mf = p->cfg.precision == AMT_PREC_MIXED_BF16;
  if (r_prev)
    launch_mg0_pre_upd(p, p->op_upd, r_prev, (double*)r, q, p->u, p_new, p->fcg,
                       (int)p->op_apply.nb, e0, e1);
  else if (mf)
    hipLaunchKernelGGL(k_mg0_pre_mfma, dim3(p->nb_tile), dim3(kBlock), 0, p->stream, r, p->g,
                       p->tg, p->s32, p->Kh16, p->Kl16, p->fixcnt, p->scal + S_P, mg[0].wd,
                       mg[0].x, mg[0].t, done);
  else
    launch_mg0_pre(p, p->op_post, r, done);
  // Stencil levels 1 .. L-2, then the dense level L-1. Levels below f run one launch per step
  // (each restriction also writes the level's first smoothing step from zero); levels from f
  // on the fused brick kernels (k_mg_down / k_mg_up) and the dense level k_mg_rdense. Levels
  // with buffers for it (l >= 2, mg_coarse_cheb > 0) smooth with the Chebyshev polynomial:
  // pre from zero, then post from the corrected iterate, k steps each (kChebCoarseExtra more on
  // the two stencil levels above the dense one on large grids), the residual for the restriction
  // formed after the last pre step.
  const int nl = (int)L, k_ch = p->cfg.mg_coarse_cheb;
  const bool cheb_extra = p->n_e > kChebExtraMinElements;
  const bool fused = p->mg_fuse_from > 0;
  const int f = fused ? p->mg_fuse_from : nl - 1;
  auto bgrid = [](const MgBrick& b) { return dim3(8u * (unsigned)((b.nbi * b.nbj * b.nbk + 7) / 8)); };
  auto plane_grid = [](const MgGrid& g) { return dim3(mg_plane_grid(g.n_nodes)); };
  // Chebyshev steps s0 .. k on level l from buffers (xi, di); returns the buffer of the result
  auto cheb_steps = [&](int l, int s0, int xi, int di) {
    const int kd = k_ch + (cheb_extra && l >= nl - 3 ? kChebCoarseExtra : 0);
    for (int st = s0; st <= kd; ++st) {
      int xo = 0, dn = -1;
      while (xo == xi || xo == di) ++xo;
      if (st < kd) {
        dn = 0;
        while (dn == xi || dn == di || dn == xo) ++dn;
      }
      const ChebCoef c = cheb_coef(st);
      hipLaunchKernelGGL(k_mg_cheb, plane_grid(mg[l].g), dim3(3 * kPlaneNodes), 0, p->stream,
                         mg[l].A, mg[l].g, mg[l].mi, mg[l].rhs, mg[l].cx[xi], mg[l].cx[di],
                         (float)c.a, (float)c.b, mg[l].cx[xo], dn >= 0 ? mg[l].cx[dn] : nullptr);
      xi = xo;
      di = dn;
    }
    return xi;
  };
  int xpre[32] = {};  // Chebyshev levels: buffer of the pre-smoothed iterate
  for (int l = 1; l <= nl - 2; ++l) {
    const bool fz = fused && l >= f, ch = mg[l].mi != nullptr;
    if (!fz) {
      hipLaunchKernelGGL(k_mg_restrict, dim3(nblk(mg[l].g.n_nodes)), dim3(kBlock), 0, p->stream,
                         mg[l - 1].t, mg[l - 1].g, mg[l].rhs, mg[l].g, mg[l].wd,
                         ch ? mg[l].cx[0] : mg[l].x);
      if (!ch) {
        if (mg[l].Ah)
          hipLaunchKernelGGL(k_mg_pre_h, dim3(mg_row_grid(mg[l].g.n_nodes)), dim3(kBlock), 0,
                             p->stream, mg[l].Ah, mg[l].Ahs, mg[l].g, mg[l].rhs, mg[l].x,
                             mg[l].t);
        else
          hipLaunchKernelGGL(k_mg_pre, dim3(mg_row_grid(mg[l].g.n_nodes)), dim3(kBlock), 0,
                             p->stream, mg[l].A, mg[l].g, mg[l].rhs, mg[l].x, mg[l].t);
        continue;
      }
      xpre[l] = cheb_steps(l, 2, 0, 0);  // x1 = d1 in buffer 0
    } else if (!ch) {
      hipLaunchKernelGGL(k_mg_down, bgrid(mg[l].brick), dim3(kFusedThreads), 0, p->stream,
                         mg[l - 1].t, mg[l - 1].g, mg[l].A, mg[l].wd, mg[l].g, mg[l].brick,
                         mg[l].rhs, mg[l].t, MgCheb{});
      continue;
    } else {
      MgCheb c;
      c.mi = mg[l].mi;
      c.a = (float)cheb_coef(2).a;
      c.b = (float)cheb_coef(2).b;
      c.x_out = mg[l].cx[0];
      c.d_out = mg[l].cx[1];
      hipLaunchKernelGGL(k_mg_down, bgrid(mg[l].brick), dim3(kFusedThreads), 0, p->stream,
                         mg[l - 1].t, mg[l - 1].g, mg[l].A, mg[l].wd, mg[l].g, mg[l].brick,
                         mg[l].rhs, mg[l].t, c);
      xpre[l] = cheb_steps(l, 3, 0, 1);
    }
    // the residual after the Chebyshev pre-smoothing, restricted by the next level
    hipLaunchKernelGGL(k_mg_pre3, plane_grid(mg[l].g), dim3(3 * kPlaneNodes), 0, p->stream,
                       mg[l].A, mg[l].g, mg[l].rhs, mg[l].cx[xpre[l]], mg[l].t);
  }
  if (fused) {
    hipLaunchKernelGGL(k_mg_rdense, dim3((p->mg_nc + 15) / 16), dim3(kBlock), 0, p->stream,
                       mg[nl - 2].t, mg[nl - 2].g, mg[nl - 1].g, p->mg_ainv, p->mg_nc,
                       mg[nl - 1].x);
  } else {
    hipLaunchKernelGGL(k_mg_restrict, dim3(nblk(mg[nl - 1].g.n_nodes)), dim3(kBlock), 0,
                       p->stream, mg[nl - 2].t, mg[nl - 2].g, mg[nl - 1].rhs, mg[nl - 1].g,
                       nullptr, mg[nl - 1].x);
    hipLaunchKernelGGL(k_mg_dense, dim3((p->mg_nc + 8 * kWaves - 1) / (8 * kWaves)), dim3(kBlock),
                       0, p->stream, p->mg_ainv, p->mg_nc, mg[nl - 1].rhs, mg[nl - 1].x);
  }
  const float* res = mg[nl - 1].x;  // correction of the level below the current one
  for (int l = nl - 2; l >= 1; --l) {
    const bool fz = fused && l >= f, ch = mg[l].mi != nullptr;
    if (!ch) {
      if (fz) {
        hipLaunchKernelGGL(k_mg_up, bgrid(mg[l].brick), dim3(kFusedThreads), 0, p->stream, res,
                           mg[l + 1].g, mg[l].A, mg[l].wd, mg[l].rhs, mg[l].g, mg[l].brick,
                           mg[l].t, MgCheb{});
      } else {
        hipLaunchKernelGGL(k_mg_prolong, dim3(nblk(mg[l].g.n_nodes)), dim3(kBlock), 0, p->stream,
                           res, mg[l + 1].g, mg[l].x, mg[l].g, nullptr);
        if (mg[l].Ah)
          hipLaunchKernelGGL(k_mg_post_h, dim3(mg_row_grid(mg[l].g.n_nodes)), dim3(kBlock), 0,
                             p->stream, mg[l].Ah, mg[l].Ahs, mg[l].g, mg[l].wd, mg[l].rhs,
                             mg[l].x, mg[l].t);
        else
          hipLaunchKernelGGL(k_mg_post, dim3(mg_row_grid(mg[l].g.n_nodes)), dim3(kBlock), 0,
                             p->stream, mg[l].A, mg[l].g, mg[l].wd, mg[l].rhs, mg[l].x,
                             mg[l].t);
      }
      res = mg[l].t;
      continue;
    }
    const int xp = xpre[l];
    int out;
    if (fz) {
      int x1 = 0, d1 = 1;
      if (x1 == xp) x1 = 2;
      if (d1 == xp) d1 = 2;
      MgCheb c;
      c.mi = mg[l].mi;
      c.b = (float)cheb_coef(1).b;
      c.x_out = mg[l].cx[x1];
      c.d_out = mg[l].cx[d1];
      c.xpre = mg[l].cx[xp];
      hipLaunchKernelGGL(k_mg_up, bgrid(mg[l].brick), dim3(kFusedThreads), 0, p->stream, res,
                         mg[l + 1].g, mg[l].A, mg[l].wd, mg[l].rhs, mg[l].g, mg[l].brick,
                         mg[l].t, c);
      out = cheb_steps(l, 2, x1, d1);
    } else {
      hipLaunchKernelGGL(k_mg_prolong, dim3(nblk(mg[l].g.n_nodes)), dim3(kBlock), 0, p->stream,
                         res, mg[l + 1].g, mg[l].cx[xp], mg[l].g, nullptr);
      out = cheb_steps(l, 1, xp, xp);
    }
    res = mg[l].cx[out];
  }
  hipLaunchKernelGGL(k_mg_prolong, dim3(nblk(p->n_n)), dim3(kBlock), 0, p->stream, res,
                     mg[1].g, mg[0].x, mg[0].g, p->fixcnt);
  if (mf)
    hipLaunchKernelGGL(k_mg0_post_mfma, dim3(p->nb_tile), dim3(kBlock), 0, p->stream, r, p->g,
                       p->tg, p->s32, p->Kh16, p->Kl16, p->fixcnt, p->scal + S_P, mg[0].wd,
                       mg[0].x, z, q, p->part(R_INIT), p->stride_red, done);
  else
    launch_mg0_post(p, p->op_post, r, z, q, p->part(R_INIT), p->stride_red, done);
}

// kMgChunk flexible-CG iterations (apply, update, V-cycle, scalar). kMgChunk is even, so
// every chunk starts on the same p-buffer parity and one captured graph serves all chunks.
void fcg_enqueue_chunk(amt_problem* p, bool eager_prof) {
  const int nba = (int)p->op_apply.nb, nbu = (int)nred(p->n_dof), nbz = mg0_post_blocks(p);
  int* done_flag = (int*)(p->scal + S_FLAG);
  // FP32 level-0 smoothing, and few enough operator blocks that every block of the fused launch
  // can reduce all the operator partials itself (nba^2 doubles of L2 reads per launch)
  const bool fused = p->cfg.precision != AMT_PREC_MIXED_BF16 && p->op_apply.nb <= kFusedUpdateTiles;
  for (int t = 0; t < kMgChunk; ++t) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (eager_prof) prof_take(p, 0, &e0, &e1);
    launch_fcg_apply(p, p->op_apply, p->z, p->p[t & 1], p->p[(t + 1) & 1], p->fcg, e0, e1);
    if (fused) {
      // the update runs inside the level-0 pre-smoothing; r alternates between two buffers
      hipEvent_t u0 = nullptr, u1 = nullptr;
      if (eager_prof) prof_take(p, 1, &u0, &u1);
      mg_vcycle(p, p->rr[(t + 1) & 1], p->z, p->q, p->rr[t & 1], p->p[(t + 1) & 1], u0, u1);
    } else {
      hipLaunchKernelGGL(k_fcg_update, dim3(nbu), dim3(kBlock), 0, p->stream, p->u, p->r,
                         p->p[(t + 1) & 1], p->q, p->n_dof, p->fcg,
                         (const double*)p->part(R_APPLY), nba, p->part(R_UPD));
      mg_vcycle(p, p->r, p->z, p->q);
    }
    hipLaunchKernelGGL(k_fcg_scalar, dim3(1), dim3(kBlock), 0, p->stream, p->fcg,
                       (const double*)p->part(R_UPD), fused ? (int)p->op_upd.nb : nbu,
                       (const double*)p->part(R_APPLY), nba, (const double*)p->part(R_INIT), nbz,
                       p->stride_red, 0, done_flag);
  }
}

// Captures the chunk once per handle. Returns false if no graph could be built, in which
// case the eager launches are used.
bool fcg_graph_ready(amt_problem* p) {
  if (p->fcg_exec) return true;
  if (p->fcg_graph_failed) return false;
  p->fcg_graph_failed = true;
  if (hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  fcg_enqueue_chunk(p, false);
  hipGraph_t g = nullptr;
  const hipError_t e1 = hipStreamEndCapture(p->stream, &g);
  if (e1 == hipSuccess && g &&
      hipGraphInstantiate(&p->fcg_exec, g, nullptr, nullptr, 0) == hipSuccess) {
    p->fcg_graph = g;
    p->fcg_graph_failed = false;
    return true;
  }
  (void)hipGetLastError();
  if (g) (void)hipGraphDestroy(g);
  p->fcg_exec = nullptr;
  return false;
}

// Flexible CG (FP64) preconditioned by the V-cycle, from x = 0 or the previous displacement.
int solve_mg(amt_problem* p, amt_solve_stats* st) {
  AMT_CHECK(tune_ops(p));
  const bool warm = p->cfg.warm_start && p->have_u;
  if (!warm)
    hipLaunchKernelGGL(k_fill<double>, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream,
                       p->u, 0.0, p->n_dof);
  dot_dev(p, p->Fb, p->Fb, p->scal + S_BB);
  AMT_CHECK(residual64(p, p->u, p->r, nullptr));
  if (warm) {
    // the start residual of a warm solve is a training sample while collection is armed; with a
    // network attached it also takes the learned step u += gamma M(r)
    if ((int)p->bank.size() < p->warm_collect) AMT_CHECK(bank_push(p, p->r));
    if (p->net) {
      AMT_CHECK(learned_warm_start(p));
    } else if (p->cfg.warm_start == 2 && p->u2) {
      // energy-optimal start in span{u_prev, u_prev2}: w = K u2 into q
      launch_apply64(p, p->op_apply, p->u2, p->q, p->scal + S_P);
      const unsigned nb = nred(p->n_dof);
      hipLaunchKernelGGL(k_proj_dots, dim3(nb), dim3(kBlock), 0, p->stream, p->u, p->u2, p->Fb,
                         p->r, p->q, p->n_dof, p->part(R_OC), p->stride_red);
      hipLaunchKernelGGL(k_proj_coef, dim3(1), dim3(kBlock), 0, p->stream,
                         (const double*)p->part(R_OC), (int)nb, p->stride_red, p->scal + S_PROJ);
      hipLaunchKernelGGL(k_proj_apply, dim3(nb), dim3(kBlock), 0, p->stream, p->u, p->u2, p->Fb,
                         p->r, p->q, p->n_dof, (const double*)(p->scal + S_PROJ),
                         p->part(R_MISC), p->stride_red);
      hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kBlock), 0, p->stream, p->part(R_MISC),
                         (int)nb, p->stride_red, p->scal + S_RR);
      AMT_CHECK(check_launch(p, "two-solution warm start"));
    }
  }
  hipLaunchKernelGGL(k_fcg_start, dim3(1), dim3(64), 0, p->stream, p->fcg, p->scal + S_RR,
                     p->scal + S_BB, p->cfg.pcg_tol, p->cfg.pcg_maxit);
  hipLaunchKernelGGL(k_fill<double>, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream, p->p[0],
                     0.0, p->n_dof);
  const int nba = (int)p->op_apply.nb, nbu = (int)nred(p->n_dof);
  int* done_flag = (int*)(p->scal + S_FLAG);
  mg_vcycle(p, p->r, p->z, nullptr);  // z0 = M r0
  hipLaunchKernelGGL(k_fcg_scalar, dim3(1), dim3(kBlock), 0, p->stream, p->fcg,
                     (const double*)p->part(R_UPD), nbu, (const double*)p->part(R_APPLY), nba,
                     (const double*)p->part(R_INIT), mg0_post_blocks(p), p->stride_red, 1,
                     done_flag);
  AMT_CHECK(check_launch(p, "fcg init"));
  // no host read here: the first batch of chunks is enqueued behind the set-up at once (its
  // kernels return at once if the start is already converged)
  FcgState hs;
  std::memset(&hs, 0, sizeof(hs));
  const bool graphs = p->cfg.use_graphs && fcg_graph_ready(p);
  // With profiling on, the first chunk of each solve is launched eagerly with an event pair
  // around every operator launch (a live sample inside the timed run); the rest replay the
  // captured graph.
  // The first batch runs as many chunks as the previous solve needed (minus one), so a
  // warm-started solve usually finishes after one or two host polls; kernels launched past
  // convergence return at once.
  bool first = true;
  while (!hs.done) {
    const int chunks = first ? std::max(1, p->mg_last_iters / kMgChunk - 1) : 1;
    for (int c = 0; c < chunks; ++c) {
      if (graphs && !(p->prof_on && first && c == 0)) {
        HIP_TRY(p, hipGraphLaunch(p->fcg_exec, p->stream));
      } else {
        fcg_enqueue_chunk(p, true);
      }
    }
    first = false;
    AMT_CHECK(check_launch(p, "fcg chunk"));
    AMT_CHECK(copy_out(p, &hs, p->fcg, sizeof(FcgState)));
  }
  p->mg_last_iters = hs.it;
  prof_harvest(p);
  st->iterations = hs.it;
  st->converged = hs.converged;
  st->outer_iterations = 1;
  st->rel_residual = hs.rhs2 > 0 ? std::sqrt(hs.rr / hs.rhs2) : 0.0;
  if (p->defer_true_rr) {  // |b - K u|^2 stays in scal[S_RR] for run_iteration's final read
    AMT_CHECK(residual64(p, p->u, p->q, nullptr));
    p->pending_rhs2 = hs.rhs2;
    st->true_rel_residual = -1.0;
    return AMT_OK;
  }
  double rr_true = 0.0;
  AMT_CHECK(residual64(p, p->u, p->q, &rr_true));
  st->true_rel_residual = hs.rhs2 > 0 ? std::sqrt(rr_true / hs.rhs2) : 0.0;
  return AMT_OK;
}

// out = a.b (device scalar), fixed-order reduction.
void dot_dev(amt_problem* p, const double* a, const double* b, double* out) {
  hipLaunchKernelGGL(k_dot, dim3(nred(p->n_dof)), dim3(kBlock), 0, p->stream, a, b, p->n_dof,
                     p->part(R_MISC), p->stride_red);
  hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kBlock), 0, p->stream, p->part(R_MISC),
                     (int)nred(p->n_dof), p->stride_red, out);
}

// Two events of the pool for a timed launch of the given kind (0 operator, 1 level-0
// pre-smoothing with the update) while profiling is on and the pool lasts, else null; the
// launch goes through hipExtLaunchKernelGGL, which timestamps the kernel's start and end.
void prof_take(amt_problem* p, int kind, hipEvent_t* e0, hipEvent_t* e1) {
  *e0 = *e1 = nullptr;
  if (!p->prof_on || p->ev_used + 2 > (int)p->ev_pool.size()) return;
  *e0 = p->ev_pool[p->ev_used];
  *e1 = p->ev_pool[p->ev_used + 1];
  p->ev_kind[p->ev_used / 2] = kind;
  p->ev_used += 2;
}
// Events for the next operator launch when it is sampled (every prof_every-th), else null.
void prof_pick(amt_problem* p, hipEvent_t* e0, hipEvent_t* e1) {
  *e0 = *e1 = nullptr;
  if (!p->prof_on) return;
  if (++p->prof_tick % p->prof_every) return;
  prof_take(p, 0, e0, e1);
}
void prof_harvest(amt_problem* p) {
  if (!p->prof_on || p->ev_used == 0) return;
  (void)hipStreamSynchronize(p->stream);
  for (int i = 0; i < p->ev_used; i += 2) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, p->ev_pool[i], p->ev_pool[i + 1]) == hipSuccess) {
      if (p->ev_kind[i / 2] == 1) {
        p->prof_ms_upd += ms;
        p->prof_launches_upd += 1;
      } else {
        p->prof_ms += ms;
        p->prof_launches += 1;
      }
    }
  }
  p->ev_used = 0;
}

// PCG from the current r (x already holds the start vector): init + start, then chunks of
// (apply, update) pairs with one device->host poll of the check flag per chunk.
template <typename T, typename TK, typename TX>
int pcg_run(amt_problem* p, const double* b, TX* x, T* r, T* z, T** pp, T* q, const T* s,
            const TK* Ke, const T* dinv, double tol, int maxit, PcgState* out,
            const std::function<bool()>* on_chunk = nullptr, bool mfma = false) {
  const GridDims g = p->g;
  const int nbu = (int)nred(p->n_dof), nba = (int)p->nb_tile;
  hipLaunchKernelGGL(k_fill<T>, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream, pp[0], T(0),
                     p->n_dof);
  hipLaunchKernelGGL((k_pcg_init<T>), dim3(nbu), dim3(kBlock), 0, p->stream, b, r, z, dinv,
                     p->n_dof, p->part(R_INIT), p->stride_red);
  hipLaunchKernelGGL(k_pcg_start, dim3(1), dim3(kBlock), 0, p->stream,
                     (const double*)p->part(R_INIT), nbu, p->stride_red, b ? 1 : 0, tol, maxit,
                     p->pcg);
  AMT_CHECK(check_launch(p, "pcg init"));
  HIP_TRY(p, hipMemcpyAsync(out, p->pcg, sizeof(PcgState), hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(p, hipStreamSynchronize(p->stream));
  int t = 0;
  while (!out->done) {
    for (int c = 0; c < kChunk; ++c, ++t) {
      // a sampled launch (events e0, e1) is timestamped by the GPU at the kernel's start and end
      hipEvent_t e0, e1;
      prof_pick(p, &e0, &e1);
      if constexpr (std::is_same<T, float>::value) {
        if (mfma)
          hipExtLaunchKernelGGL(k_pcg_apply_mfma, dim3(nba), dim3(kBlock), 0, p->stream, e0, e1,
                                0, t, z, pp[t & 1], pp[(t + 1) & 1], q, s, p->Kh16, p->Kl16,
                                p->fixcnt, p->scal + S_P, g, p->tg, p->pcg,
                                (const double*)p->part(R_UPD), nbu, p->part(R_APPLY),
                                p->stride_red);
        else
          hipExtLaunchKernelGGL((k_pcg_apply<T, TK>), dim3(nba), dim3(kBlock), 0, p->stream, e0,
                                e1, 0, t, z, pp[t & 1], pp[(t + 1) & 1], q, s, Ke, p->fixcnt,
                                p->scal + S_P, g, p->tg, p->pcg, (const double*)p->part(R_UPD),
                                nbu, p->part(R_APPLY), p->stride_red);
      } else {
        hipExtLaunchKernelGGL((k_pcg_apply<T, TK>), dim3(nba), dim3(kBlock), 0, p->stream, e0, e1,
                              0, t, z, pp[t & 1], pp[(t + 1) & 1], q, s, Ke, p->fixcnt,
                              p->scal + S_P, g, p->tg, p->pcg, (const double*)p->part(R_UPD), nbu,
                              p->part(R_APPLY), p->stride_red);
      }
      hipLaunchKernelGGL((k_pcg_update<T, TX>), dim3(nbu), dim3(kBlock), 0, p->stream, t, x, r,
                         z, pp[(t + 1) & 1], q, dinv, p->n_dof, p->pcg,
                         (const double*)p->part(R_APPLY), nba, p->part(R_UPD), p->stride_red);
    }
    hipLaunchKernelGGL(k_pcg_check, dim3(1), dim3(kBlock), 0, p->stream, t, p->pcg,
                       (const double*)p->part(R_UPD), nbu, p->stride_red, p->pcg_flag);
    AMT_CHECK(check_launch(p, "pcg chunk"));
    HIP_TRY(p, hipMemcpyAsync(out, p->pcg_flag, sizeof(PcgState), hipMemcpyDeviceToHost,
                              p->stream));
    HIP_TRY(p, hipStreamSynchronize(p->stream));
    if (on_chunk && !out->done && !(*on_chunk)()) break;
  }
  prof_harvest(p);
  return AMT_OK;
}

// FP64 residual r_out = Fb - K x and |r|^2 into scal[S_RR] (blocking read).
int residual64(amt_problem* p, const double* x, double* rout, double* rr_host) {
  const int nb = (int)launch_residual(p, x, rout);
  hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kBlock), 0, p->stream, p->part(R_APPLY), nb,
                     nb, p->scal + S_RR);
  AMT_CHECK(check_launch(p, "k_residual"));
  if (rr_host) {
    HIP_TRY(p, hipMemcpyAsync(rr_host, p->scal + S_RR, sizeof(double), hipMemcpyDeviceToHost,
                              p->stream));
    HIP_TRY(p, hipStreamSynchronize(p->stream));
  }
  return AMT_OK;
}

__global__ void k_cast_scale(const double* __restrict__ a, float* __restrict__ b, double sc,
                             int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) b[i] = (float)(a[i] * sc);
}
__global__ void k_axpy_cast(double* __restrict__ x, const float* __restrict__ d, double sc,
                            int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = fma((double)d[i], sc, x[i]);
}

// Plain FP64 Jacobi-PCG (Eigen loop), from x = 0 or from the previous displacement.
int solve_fp64(amt_problem* p, amt_solve_stats* st) {
  const bool warm = p->cfg.warm_start && p->have_u;
  if (warm) {
    AMT_CHECK(residual64(p, p->u, p->r, nullptr));
  } else {
    hipLaunchKernelGGL(k_fill<double>, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream, p->u,
                       0.0, p->n_dof);
    HIP_TRY(p, hipMemcpyAsync(p->r, p->Fb, sizeof(double) * p->n_dof, hipMemcpyDeviceToDevice,
                              p->stream));
  }
  PcgState hs;
  AMT_CHECK((pcg_run<double, double, double>(p, p->Fb, p->u, p->r, p->z, p->p, p->q, p->s64,
                                             p->Ke64, p->dinv, p->cfg.pcg_tol,
                                             p->cfg.pcg_maxit, &hs)));
  if (hs.rhs2 == 0.0)  // Eigen: zero rhs -> x = 0
    hipLaunchKernelGGL(k_fill<double>, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream,
                       p->u, 0.0, p->n_dof);
  double rr_true = 0.0;
  AMT_CHECK(residual64(p, p->u, p->q, &rr_true));
  st->iterations = hs.it;
  st->converged = hs.converged;
  st->outer_iterations = 1;
  st->fallback = 0;
  st->rel_residual = hs.rhs2 > 0 ? std::sqrt(hs.rr / hs.rhs2) : 0.0;
  st->true_rel_residual = hs.rhs2 > 0 ? std::sqrt(rr_true / hs.rhs2) : 0.0;
  return AMT_OK;
}

// FP64 iterative refinement around FP32 Jacobi-PCG inner solves:
//   r = F_bc - K x (FP64);  solve K d = r/|r| in FP32 to inner_tol;  x += |r| d.
// The stopping test is the FP64 true residual |r| <= pcg_tol |F_bc|.
int solve_mixed(amt_problem* p, amt_solve_stats* st) {
  const bool warm = p->cfg.warm_start && p->have_u;
  if (!warm)
    hipLaunchKernelGGL(k_fill<double>, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream,
                       p->u, 0.0, p->n_dof);
  dot_dev(p, p->Fb, p->Fb, p->scal + S_BB);
  double bb = 0.0, rr = 0.0;
  HIP_TRY(p, hipMemcpyAsync(&bb, p->scal + S_BB, sizeof(double), hipMemcpyDeviceToHost,
                            p->stream));
  AMT_CHECK(residual64(p, p->u, p->r, &rr));
  const double thresh = std::max(p->cfg.pcg_tol * p->cfg.pcg_tol * bb, 2.2250738585072014e-308);
  int inner_total = 0, outer = 0;
  const int max_outer = 60;
  PcgState hs;
  std::memset(&hs, 0, sizeof(hs));
  while (bb > 0.0 && rr >= thresh && outer < max_outer && inner_total < p->cfg.pcg_maxit) {
    const double nr = std::sqrt(rr);
    hipLaunchKernelGGL(k_cast_scale, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream, p->r,
                       p->r32, 1.0 / nr, p->n_dof);
    hipLaunchKernelGGL(k_fill<float>, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream, p->d32,
                       0.0f, p->n_dof);
    const int budget = std::min(p->cfg.inner_maxit, p->cfg.pcg_maxit - inner_total);
    AMT_CHECK((pcg_run<float, float, float>(p, nullptr, p->d32, p->r32, p->z32, p->p32,
                                            p->q32, p->s32, p->Ke32, p->dinv32,
                                            p->cfg.inner_tol, budget, &hs, nullptr,
                                            p->cfg.precision == AMT_PREC_MIXED_BF16)));
    inner_total += std::max(hs.it, 1);
    hipLaunchKernelGGL(k_axpy_cast, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream, p->u,
                       p->d32, nr, p->n_dof);
    AMT_CHECK(residual64(p, p->u, p->r, &rr));
    ++outer;
  }
  st->iterations = inner_total;
  st->outer_iterations = outer;
  st->converged = (bb == 0.0) || (rr < thresh);
  st->rel_residual = bb > 0 ? std::sqrt(rr / bb) : 0.0;
  st->true_rel_residual = st->rel_residual;
  return AMT_OK;
}


// ---- learned preconditioner: a BF16 U-Net correction on top of Jacobi ----

void nn_dims(const amt_problem* p, int& px, int& py, int& pz) {
  px = (p->g.nx + 1 + 7) / 8 * 8;
  py = (p->g.ny + 1 + 7) / 8 * 8;
  pz = (p->g.nz + 1 + 7) / 8 * 8;
}

int nn_check(amt_problem* p, amt_unet* net) {
  if (!net) return fail(p, AMT_EINVAL, "no U-Net attached");
  int px, py, pz;
  nn_dims(p, px, py, pz);
  if (net->d[0].px != px || net->d[0].py != py || net->d[0].pz != pz)
    return fail(p, AMT_EINVAL, "U-Net grid does not match the problem (use amt_unet_create_for)");
  return AMT_OK;
}

int nn_alloc(amt_problem* p) {
  if (p->nn_feat) return AMT_OK;
  AMT_CHECK(dalloc(p, &p->nn_feat, p->n_n));
  AMT_CHECK(dalloc(p, &p->nn_U, p->n_dof));
  AMT_CHECK(dalloc(p, &p->nn_AU, p->n_dof));
  AMT_CHECK(dalloc(p, &p->nn_J, p->n_dof));
  AMT_CHECK(dalloc(p, &p->nn_KJ, p->n_dof));
  return AMT_OK;
}

// Network forward on residual r: X0 from the density feature and D^-1 r / rms.
// infer: the forward pass without the training intermediates (no backward follows).
template <typename T>
void nn_forward(amt_problem* p, amt_unet* net, const T* r, bool infer) {
  net->stream = p->stream;
  const unsigned nbz = nred(p->n_dof);
  hipLaunchKernelGGL(k_nn_zz<T>, dim3(nbz), dim3(kBlock), 0, p->stream, r, p->dinv, p->n_dof,
                     p->part(R_MISC));
  hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kBlock), 0, p->stream, p->part(R_MISC),
                     (int)nbz, (int)nbz, p->scal + S_NNZZ);
  const VoxDims& d = net->d[0];
  hipLaunchKernelGGL(k_nn_pack<T>, dim3(nblk(d.nv * 16)), dim3(kBlock), 0, p->stream, r, p->dinv,
                     p->nn_feat, p->g, d, p->scal + S_NNZZ, net->act[amt_unet::X0]);
  amt_unet_impl::forward(net, infer);
}

// z = M(r) = D^-1 r + gain * rms * net(feat, D^-1 r / rms); partials of z.r, z.q into R_INIT
// (nred(n_dof) of them).
template <typename T>
void nn_apply(amt_problem* p, amt_unet* net, const T* r, double* z64, float* z32,
              const double* q) {
  nn_forward<T>(p, net, r, true);
  hipLaunchKernelGGL(k_nn_unpack<T>, dim3(nred(p->n_dof)), dim3(kBlock), 0, p->stream,
                     net->act[amt_unet::OUT], net->d[0], p->g, r, p->dinv, p->diag, p->fixcnt,
                     p->scal + S_NNZZ, net->gain, z64, z32, q, p->part(R_INIT));
}

// Flexible CG (FP64) with the learned preconditioner. Stagnation (kNnPatience chunks without
// a 10% drop of |r| below its last mark), a non-finite residual or a non-positive curvature
// hands the solve to plain FP64 Jacobi-PCG from the current iterate; st->fallback reports it.
int solve_unet(amt_problem* p, amt_solve_stats* st) {
  amt_unet* net = p->net;
  AMT_CHECK(nn_check(p, net));
  const bool warm = p->cfg.warm_start && p->have_u;
  dot_dev(p, p->Fb, p->Fb, p->scal + S_BB);
  if (!warm) {
    // learned warm start: u0 = gamma M(F_bc), gamma the energy-optimal multiple
    nn_apply<double>(p, net, p->Fb, p->z, nullptr, nullptr);
    launch_apply64(p, p->op_apply, p->z, p->q, (const double*)(p->scal + S_P));
    dot_dev(p, p->z, p->Fb, p->scal + S_NNUAU);
    dot_dev(p, p->z, p->q, p->scal + S_NNUR);
    hipLaunchKernelGGL(k_nn_scale, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream, p->u, p->z,
                       p->scal + S_NNUAU, p->n_dof);
  }
  AMT_CHECK(residual64(p, p->u, p->r, nullptr));
  hipLaunchKernelGGL(k_fcg_start, dim3(1), dim3(64), 0, p->stream, p->fcg, p->scal + S_RR,
                     p->scal + S_BB, p->cfg.pcg_tol, p->cfg.pcg_maxit);
  hipLaunchKernelGGL(k_fill<double>, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream, p->p[0],
                     0.0, p->n_dof);
  // k_nn_unpack runs nred(n_dof) blocks: its partials fit the R_INIT region
  const int nba = (int)p->op_apply.nb, nbu = (int)nred(p->n_dof), nbz = nbu;
  int* done_flag = (int*)(p->scal + S_FLAG);
  nn_apply<double>(p, net, p->r, p->z, nullptr, nullptr);
  hipLaunchKernelGGL(k_fcg_scalar, dim3(1), dim3(kBlock), 0, p->stream, p->fcg,
                     (const double*)p->part(R_UPD), nbu, (const double*)p->part(R_APPLY), nba,
                     (const double*)p->part(R_INIT), nbz, nbz, 1, done_flag);
  AMT_CHECK(check_launch(p, "unet fcg init"));
  FcgState hs;
  AMT_CHECK(copy_out(p, &hs, p->fcg, sizeof(FcgState)));
  // Stagnation test: |r| must drop 10% below its last mark within kNnPatience chunks.
  double rr_mark = hs.rr;
  int t = 0, fell_back = 0, chunk = 0, mark_chunk = 0;
  while (!hs.done) {
    for (int c = 0; c < kNnChunk; ++c, ++t) {
      launch_fcg_apply(p, p->op_apply, p->z, p->p[t & 1], p->p[(t + 1) & 1], p->fcg);
      hipLaunchKernelGGL(k_fcg_update, dim3(nbu), dim3(kBlock), 0, p->stream, p->u, p->r,
                         p->p[(t + 1) & 1], p->q, p->n_dof, p->fcg,
                         (const double*)p->part(R_APPLY), nba, p->part(R_UPD));
      nn_apply<double>(p, net, p->r, p->z, nullptr, p->q);
      hipLaunchKernelGGL(k_fcg_scalar, dim3(1), dim3(kBlock), 0, p->stream, p->fcg,
                         (const double*)p->part(R_UPD), nbu, (const double*)p->part(R_APPLY),
                         nba, (const double*)p->part(R_INIT), nbz, nbz, 0, done_flag);
    }
    AMT_CHECK(check_launch(p, "unet fcg chunk"));
    AMT_CHECK(copy_out(p, &hs, p->fcg, sizeof(FcgState)));
    if (hs.done) break;
    ++chunk;
    if (!std::isfinite(hs.rr) || !(hs.pq > 0.0)) {
      fell_back = 1;
      break;
    }
    if (hs.rr < 0.81 * rr_mark) {
      rr_mark = hs.rr;
      mark_chunk = chunk;
    } else if (chunk - mark_chunk >= kNnPatience) {
      fell_back = 1;
      break;
    }
  }
  if (fell_back) {
    const int used = hs.it;
    const int saved_warm = p->cfg.warm_start, saved_maxit = p->cfg.pcg_maxit;
    p->cfg.warm_start = 1;
    p->have_u = true;
    p->cfg.pcg_maxit = std::max(1, saved_maxit - used);
    const int rc = solve_fp64(p, st);
    p->cfg.warm_start = saved_warm;
    p->cfg.pcg_maxit = saved_maxit;
    AMT_CHECK(rc);
    st->iterations += used;
    st->outer_iterations = 2;
    st->fallback = 1;
    return AMT_OK;
  }
  double rr_true = 0.0;
  AMT_CHECK(residual64(p, p->u, p->q, &rr_true));
  st->iterations = hs.it;
  st->converged = hs.converged;
  st->outer_iterations = 1;
  st->fallback = 0;
  st->rel_residual = hs.rhs2 > 0 ? std::sqrt(hs.rr / hs.rhs2) : 0.0;
  st->true_rel_residual = hs.rhs2 > 0 ? std::sqrt(rr_true / hs.rhs2) : 0.0;
  return AMT_OK;
}

// Learned warm start of a multigrid solve: with r = F_bc - K u (from residual64), z = M(r) (the
// network's correction on top of Jacobi), u += gamma z with gamma = z.r / z.K z (the step that
// minimises the energy error along z; 0 if z.K z is not positive), then r and |r|^2 afresh in
// FP64. Two operator applications and one network pass per solve.
int learned_warm_start(amt_problem* p) {
  amt_unet* net = p->net;
  AMT_CHECK(nn_check(p, net));
  if (!p->nn_feat) {
    AMT_CHECK(nn_alloc(p));
    hipLaunchKernelGGL(k_nn_feature, dim3(nblk(p->n_n)), dim3(kBlock), 0, p->stream, p->s64, p->g,
                       p->nn_feat);
  }
  net->stream = p->stream;
  nn_apply<double>(p, net, p->r, p->z, nullptr, nullptr);
  launch_apply64(p, p->op_apply, p->z, p->q, (const double*)(p->scal + S_P));
  dot_dev(p, p->z, p->r, p->scal + S_NNUAU);
  dot_dev(p, p->z, p->q, p->scal + S_NNUR);
  hipLaunchKernelGGL(k_nn_step, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream, p->u, p->z,
                     (const double*)(p->scal + S_NNUAU), p->n_dof);
  AMT_CHECK(check_launch(p, "learned warm start"));
  return residual64(p, p->u, p->r, nullptr);
}

// Appends a training sample: r (FP64, converted) with a copy of the current density.
int bank_push(amt_problem* p, const double* r) {
  amt_problem::Sample s{nullptr, nullptr};
  AMT_CHECK(dalloc(p, &s.r, p->n_dof));
  AMT_CHECK(dalloc(p, &s.rho, p->n_e));
  hipLaunchKernelGGL(k_d2f, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream, r, s.r, p->n_dof);
  HIP_TRY(p, hipMemcpyAsync(s.rho, p->rho, sizeof(double) * p->n_e, hipMemcpyDeviceToDevice,
                            p->stream));
  p->bank.push_back(s);
  return AMT_OK;
}

// The coefficients a training step reads, for a sample's density: element scales, diagonal,
// penalty, Jacobi inverse and the network's density feature (no multigrid set-up).
void sample_coeffs(amt_problem* p, const double* rho) {
  (void)hipMemcpyAsync(p->rho, rho, sizeof(double) * p->n_e, hipMemcpyDeviceToDevice,
                       p->stream);
  const GridDims g = p->g;
  hipLaunchKernelGGL(k_simp, dim3(nblk(p->n_e)), dim3(kBlock), 0, p->stream, p->rho, p->s64,
                     p->s32, g, p->cfg.density_index == AMT_INDEX_REF_COMPAT ? 1 : 0,
                     p->cfg.penalty, p->cfg.E_min, p->cfg.rho_min);
  hipLaunchKernelGGL(k_diag, dim3(nblk(p->n_n)), dim3(kBlock), 0, p->stream, p->diag, p->s64,
                     p->Ke64, g, p->part(R_APPLY), p->stride_apply);
  hipLaunchKernelGGL((k_finalize<1, true>), dim3(1), dim3(kBlock), 0, p->stream,
                     p->part(R_APPLY), (int)nblk(p->n_n), p->stride_apply, p->scal + S_MAXDIAG);
  hipLaunchKernelGGL(k_dinv, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream, p->diag,
                     p->fixcnt, p->scal + S_MAXDIAG, p->cfg.penalty_factor, p->scal + S_P,
                     p->dinv, p->dinv32, p->n_dof);
  hipLaunchKernelGGL(k_nn_feature, dim3(nblk(p->n_n)), dim3(kBlock), 0, p->stream, p->s64, p->g,
                     p->nn_feat);
}

// One training step on residual sample r (FP32): energy of U = J + theta N (J = D^-1 r, N the
// network's correction, theta from the running means), two FP32 operator applications;
// terms (optional) gets {U.KU, U.r, r.D^-1 r, J.KJ, J.r}. lr <= 0 only evaluates.
void nn_train_step(amt_problem* p, amt_unet* net, const float* r, float lr, double* terms) {
  const unsigned nbz = nred(p->n_dof);
  hipLaunchKernelGGL(k_nn_rdr, dim3(nbz), dim3(kBlock), 0, p->stream, r, p->dinv, p->n_dof,
                     p->part(R_UPD));
  hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kBlock), 0, p->stream, p->part(R_UPD),
                     (int)nbz, (int)nbz, p->scal + S_NNRDR);
  nn_forward<float>(p, net, r, false);
  const VoxDims& d = net->d[0];
  hipLaunchKernelGGL(k_nn_fields, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream,
                     net->act[amt_unet::OUT], d, p->g, r, p->dinv, p->diag, p->fixcnt,
                     p->scal + S_NNZZ, p->nn_J, p->nn_U);
  const double* P = p->scal + S_P;
  hipLaunchKernelGGL((k_apply<float, float>), dim3(p->nb_tile), dim3(kBlock), 0, p->stream,
                     p->nn_J, p->nn_KJ, p->s32, p->Ke32, p->fixcnt, P, p->g, p->tg);
  hipLaunchKernelGGL((k_apply<float, float>), dim3(p->nb_tile), dim3(kBlock), 0, p->stream,
                     p->nn_U, p->nn_AU, p->s32, p->Ke32, p->fixcnt, P, p->g, p->tg);
  // 5 partials per block: R_OC (idle during training) holds kOcCand * stride_red doubles;
  // the running means' ratio is the weight of N in this step's U and its gradient
  hipLaunchKernelGGL(k_nn_dots, dim3(nbz), dim3(kBlock), 0, p->stream, r, p->nn_J, p->nn_KJ,
                     p->nn_U, p->nn_AU, p->n_dof, p->part(R_OC), (int)nbz);
  hipLaunchKernelGGL((k_finalize<5>), dim3(1), dim3(kBlock), 0, p->stream, p->part(R_OC),
                     (int)nbz, (int)nbz, p->scal + S_NND0);
  const bool eval_only = !(lr > 0.f);
  hipLaunchKernelGGL(k_nn_theta, dim3(1), dim3(64), 0, p->stream, p->scal + S_NND0,
                     p->scal + S_NNRDR, net->gain, eval_only ? 0.0 : kNnGainMomentum,
                     p->scal + S_NNTH, terms);
  if (eval_only) return;
  hipLaunchKernelGGL(k_nn_loss_grad, dim3(nblk(d.nv * 16)), dim3(kBlock), 0, p->stream,
                     p->nn_KJ, p->nn_AU, r, p->diag, p->fixcnt, p->g, d, p->scal + S_NNZZ,
                     p->scal + S_NNTH, p->scal + S_NNRDR, net->grd[amt_unet::OUT]);
  hipLaunchKernelGGL(k_fill<float>, dim3(nblk(net->n_params)), dim3(kBlock), 0, p->stream,
                     net->grad, 0.f, net->n_params);
  amt_unet_impl::backward(net);
  amt_unet_impl::adam_step(net, lr);
}

int do_solve(amt_problem* p, amt_solve_stats* st) {
  if (!p->loads_set) return fail(p, AMT_EINVAL, "loads not set (amt_set_loads)");
  if (!p->coeffs_ready) AMT_CHECK(prepare_coeffs(p));
  int rc;
  if (p->cfg.precond == AMT_PRECOND_MULTIGRID)
    rc = solve_mg(p, st);
  else if (p->cfg.precond == AMT_PRECOND_UNET)
    rc = solve_unet(p, st);
  else
    rc = p->cfg.precision == AMT_PREC_FP64 ? solve_fp64(p, st) : solve_mixed(p, st);
  if (rc == AMT_OK) p->have_u = true;
  return rc;
}

int build_design(amt_problem* p) {
  hipLaunchKernelGGL(k_design_mask, dim3(nred(p->n_e)), dim3(kBlock), 0, p->stream, p->rho,
                     p->design, p->n_e, p->cfg.rho_min, p->part(R_MISC), p->stride_red);
  hipLaunchKernelGGL((k_finalize<1>), dim3(1), dim3(kBlock), 0, p->stream, p->part(R_MISC),
                     (int)nred(p->n_e), p->stride_red, p->scal + S_NDESIGN);
  AMT_CHECK(check_launch(p, "k_design_mask"));
  HIP_TRY(p, hipMemcpyAsync(&p->n_design, p->scal + S_NDESIGN, sizeof(double),
                            hipMemcpyDeviceToHost, p->stream));
  HIP_TRY(p, hipStreamSynchronize(p->stream));
  p->design_ready = true;
  return AMT_OK;
}

// The OC update's launches on the device density with p->dcf (state into p->oc); every
// argument is fixed for the handle, so the sequence can be captured once.
void enqueue_oc(amt_problem* p) {
  hipLaunchKernelGGL(k_oc_init, dim3(1), dim3(64), 0, p->stream, p->oc,
                     p->cfg.volume_fraction, (const double*)(p->scal + S_NDESIGN), p->n_e);
  const unsigned nb = std::min<unsigned>(nred(p->n_e), kOcBlocks);
  hipLaunchKernelGGL(k_oc_bounds, dim3(nb), dim3(kBlock), 0, p->stream, p->rho, p->dcf,
                     p->design, p->n_e, p->cfg.move_limit, p->cfg.rho_min, p->part(R_OC),
                     p->stride_red);
  hipLaunchKernelGGL(k_oc_bounds_final, dim3(1), dim3(kBlock), 0, p->stream, p->oc,
                     (const double*)p->part(R_OC), (int)nb, p->stride_red);
  // ceil(100 / LEV) + 1 sweeps always suffice: each one advances LEV passes or finishes
  for (int sw = 0; sw < (100 + kOcLevels - 1) / kOcLevels + 1; ++sw) {
    hipLaunchKernelGGL((k_oc_sweep<kOcLevels>), dim3(nb), dim3(kBlock), 0, p->stream, p->rho,
                       p->dcf, p->design, p->n_e, p->cfg.move_limit, p->cfg.rho_min, p->oc,
                       p->part(R_OC), p->stride_red);
    hipLaunchKernelGGL((k_oc_walk<kOcLevels>), dim3(1), dim3(kBlock), 0, p->stream, p->oc,
                       (const double*)p->part(R_OC), (int)nb, p->stride_red);
  }
  hipLaunchKernelGGL(k_oc_final, dim3(nb), dim3(kBlock), 0, p->stream, p->rho, p->dcf,
                     p->design, p->n_e, p->cfg.move_limit, p->cfg.rho_min, p->oc,
                     p->part(R_MISC), p->stride_red);
  hipLaunchKernelGGL(k_oc_final_reduce, dim3(1), dim3(kBlock), 0, p->stream, p->oc,
                     (const double*)p->part(R_MISC), (int)nb, p->stride_red);
}

// OC update on the device density with p->dcf; writes state into p->oc.
int oc_update(amt_problem* p, OcState* hs) {
  if (!p->design_ready) AMT_CHECK(build_design(p));
  enqueue_oc(p);
  AMT_CHECK(check_launch(p, "oc"));
  p->coeffs_ready = false;
  if (hs) {
    HIP_TRY(p, hipMemcpyAsync(hs, p->oc, sizeof(OcState), hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(p, hipStreamSynchronize(p->stream));
  }
  return AMT_OK;
}

// After the solve: compliance with the unmodified forces (topology_optimizer.cpp:338),
// sensitivities, filter, the OC update and the end-of-iteration scalars for one host read.
void enqueue_post_solve(amt_problem* p) {
  dot_dev(p, p->u, p->F, p->scal + S_COMPL);
  hipLaunchKernelGGL(k_sens, dim3(nblk(p->n_e)), dim3(kBlock), 0, p->stream, p->u, p->rho,
                     p->Ke64, p->g, p->cfg.penalty, p->dc);
  launch_filter(p);
  enqueue_oc(p);
  hipLaunchKernelGGL(k_iter_scalars, dim3(1), dim3(64), 0, p->stream, p->oc, p->scal + S_COMPL,
                     p->scal + S_RR, p->iter_scal);
}

// The post-solve launches (~80, most of them short OC sweeps and walks) captured once per
// handle and replayed, so the host does not enqueue them one by one. Returns false if no graph
// could be built (the launches are then enqueued eagerly).
bool post_graph_ready(amt_problem* p) {
  if (p->post_exec) return true;
  if (p->post_graph_failed) return false;
  p->post_graph_failed = true;
  if (hipStreamBeginCapture(p->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  enqueue_post_solve(p);
  hipGraph_t g = nullptr;
  const hipError_t e1 = hipStreamEndCapture(p->stream, &g);
  if (e1 == hipSuccess && g &&
      hipGraphInstantiate(&p->post_exec, g, nullptr, nullptr, 0) == hipSuccess) {
    p->post_graph = g;
    p->post_graph_failed = false;
    return true;
  }
  (void)hipGetLastError();
  if (g) (void)hipGraphDestroy(g);
  p->post_exec = nullptr;
  return false;
}

int run_iteration(amt_problem* p, amt_iteration_result* out) {
  amt_iteration_result R;
  std::memset(&R, 0, sizeof(R));
  if (!p->loads_set) return fail(p, AMT_EINVAL, "loads not set (amt_set_loads)");
  if (!p->design_ready) AMT_CHECK(build_design(p));
  const double t0 = now_s();
  AMT_CHECK(prepare_coeffs(p));
  if (p->cfg.verbose) HIP_TRY(p, hipStreamSynchronize(p->stream));
  const double t1 = now_s();
  amt_solve_stats ss;
  std::memset(&ss, 0, sizeof(ss));
  p->defer_true_rr = true;
  const int rc_solve = do_solve(p, &ss);
  p->defer_true_rr = false;
  AMT_CHECK(rc_solve);
  const double t2 = now_s();
  double t3 = t2;
  if (p->cfg.use_graphs && !p->cfg.verbose && post_graph_ready(p)) {
    HIP_TRY(p, hipGraphLaunch(p->post_exec, p->stream));
  } else {
    // compliance with the unmodified forces (topology_optimizer.cpp:338)
    dot_dev(p, p->u, p->F, p->scal + S_COMPL);
    hipLaunchKernelGGL(k_sens, dim3(nblk(p->n_e)), dim3(kBlock), 0, p->stream, p->u, p->rho,
                       p->Ke64, p->g, p->cfg.penalty, p->dc);
    AMT_CHECK(check_launch(p, "sensitivities"));
    if (p->cfg.verbose) HIP_TRY(p, hipStreamSynchronize(p->stream));
    t3 = now_s();
    launch_filter(p);
    enqueue_oc(p);
    // one host read per OC iteration after the solve: OC state, compliance, true residual
    hipLaunchKernelGGL(k_iter_scalars, dim3(1), dim3(64), 0, p->stream, p->oc,
                       p->scal + S_COMPL, p->scal + S_RR, p->iter_scal);
  }
  AMT_CHECK(check_launch(p, "post-solve"));
  p->coeffs_ready = false;
  IterScalars sc;
  AMT_CHECK(copy_out(p, &sc, p->iter_scal, sizeof(IterScalars)));
  const OcState& hs = sc.oc;
  if (ss.true_rel_residual < 0.0)
    ss.true_rel_residual = p->pending_rhs2 > 0 ? std::sqrt(sc.rr / p->pending_rhs2) : 0.0;
  const double compl_ = sc.compl_;
  const double t4 = now_s();
  R.compliance = compl_;
  R.max_density_change = hs.change;
  R.bisect_passes = hs.passes;
  R.volume_fraction = p->n_design > 0 ? hs.vol_sum / p->n_design : hs.vol_sum / p->n_e;
  R.solver_iterations = ss.iterations;
  R.solver_converged = ss.converged;
  R.outer_iterations = ss.outer_iterations;
  R.solver_fallback = ss.fallback;
  R.true_rel_residual = ss.true_rel_residual;
  R.assembly_time_s = t1 - t0;
  R.solve_time_s = t2 - t1;
  R.sensitivity_time_s = t3 - t2;
  R.update_time_s = t4 - t3;
  if (!ss.converged && p->cfg.verbose)
    std::fprintf(stderr, "WARNING: Solver did not converge at iteration!\n");
  if (out) *out = R;
  return AMT_OK;
}

}  // namespace

extern "C" {

void amt_config_default(amt_config* c) {
  std::memset(c, 0, sizeof(*c));
  c->nx = c->ny = c->nz = 1;
  c->element_size = 1.0;
  c->E = 113.8e9;
  c->nu = 0.342;
  c->volume_fraction = 0.25;
  c->penalty = 3.0;
  c->filter_radius = 2.0;
  c->move_limit = 0.2;
  c->rho_min = 0.01;
  c->E_min = 1e-6;
  c->max_iterations = 100;
  c->convergence_tol = 0.01;
  c->pcg_tol = 1e-6;
  c->pcg_maxit = 5000;
  c->penalty_factor = 1e8;
  c->density_index = AMT_INDEX_REF_COMPAT;
  c->precision = AMT_PREC_FP64;
  c->precond = AMT_PRECOND_JACOBI;
  c->warm_start = 0;
  c->inner_tol = 1e-4;
  c->inner_maxit = 20000;
  c->verbose = 0;
  c->use_graphs = 1;
  c->operator_tile = -1;
  c->mg_fused = 1;
  c->mg_coarse_cheb = 3;
}

int amt_element_stiffness(double E, double nu, double h, double* Ke) {
  if (!Ke) return fail(nullptr, AMT_EINVAL, "null output");
  if (!(h > 0) || !(E > 0) || !(nu > -1.0 && nu < 0.5))
    return fail(nullptr, AMT_EINVAL, "invalid element size or material");
  hex8_stiffness(E, nu, h, Ke);
  return AMT_OK;
}

const char* amt_last_error(const amt_problem* p) {
  return p ? p->err.c_str() : g_create_error.c_str();
}

int amt_create(const amt_config* cfg, amt_problem** out) {
  if (!cfg || !out) return fail(nullptr, AMT_EINVAL, "null argument");
  *out = nullptr;
  if (cfg->nx <= 0 || cfg->ny <= 0 || cfg->nz <= 0)
    return fail(nullptr, AMT_EINVAL, "grid dimensions must be positive");
  if (!(cfg->element_size > 0) || !(cfg->E > 0) || !(cfg->nu > -1.0 && cfg->nu < 0.5))
    return fail(nullptr, AMT_EINVAL, "invalid element size or material");
  if (!(cfg->filter_radius >= 0) || cfg->filter_radius > 64)
    return fail(nullptr, AMT_EINVAL, "filter radius must be in [0, 64]");
  if (cfg->precond != AMT_PRECOND_JACOBI && cfg->precond != AMT_PRECOND_MULTIGRID &&
      cfg->precond != AMT_PRECOND_UNET)
    return fail(nullptr, AMT_EINVAL, "unknown preconditioner");
  if (cfg->precision != AMT_PREC_FP64 && cfg->precision != AMT_PREC_MIXED_FP32 &&
      cfg->precision != AMT_PREC_MIXED_BF16)
    return fail(nullptr, AMT_EINVAL, "unknown precision");
  if (cfg->warm_start < 0 || cfg->warm_start > 2)
    return fail(nullptr, AMT_EINVAL, "warm_start must be 0, 1 or 2");
  if (cfg->mg_fused != 0 && cfg->mg_fused != 1)
    return fail(nullptr, AMT_EINVAL, "mg_fused must be 0 or 1");
  if (cfg->mg_coarse_cheb != 0 && (cfg->mg_coarse_cheb < 2 || cfg->mg_coarse_cheb > 4))
    return fail(nullptr, AMT_EINVAL, "mg_coarse_cheb must be 0 or 2..4");
  if (cfg->operator_tile < -1 || cfg->operator_tile >= kOpCandidates)
    return fail(nullptr, AMT_EINVAL, "operator_tile must be -1 (autotune) or a candidate 0.." +
                                         std::to_string(kOpCandidates - 1));
  const int64_t n_n = (int64_t)(cfg->nx + 1) * (cfg->ny + 1) * (cfg->nz + 1);
  if (3 * n_n > (int64_t)INT32_MAX) return fail(nullptr, AMT_EINVAL, "grid too large");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(nullptr, AMT_EHIP, "no HIP device available");

  amt_problem* p = new amt_problem();
  p->cfg = *cfg;
  p->g.nx = cfg->nx;
  p->g.ny = cfg->ny;
  p->g.nz = cfg->nz;
  p->g.nny = cfg->ny + 1;
  p->g.nnz = cfg->nz + 1;
  p->n_e = (int64_t)cfg->nx * cfg->ny * cfg->nz;
  p->n_n = n_n;
  p->n_dof = 3 * n_n;
  p->g.n_nodes = p->n_n;
  p->g.n_elem = p->n_e;

  auto bail = [&](int rc) {
    g_create_error = p->err;
    amt_destroy(p);
    return rc;
  };
#define CREATE_TRY(x)              \
  do {                             \
    int rc_ = (x);                 \
    if (rc_ != AMT_OK) return bail(rc_); \
  } while (0)

  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(p, AMT_EHIP, "hipStreamCreate failed"));
  const int64_t nd = p->n_dof, ne = p->n_e;
  CREATE_TRY(dalloc(p, &p->Ke64, 576));
  CREATE_TRY(dalloc(p, &p->Ke32, 576));
  CREATE_TRY(dalloc(p, &p->Kh16, 576));
  CREATE_TRY(dalloc(p, &p->Kl16, 576));
  CREATE_TRY(dalloc(p, &p->rho, ne));
  CREATE_TRY(dalloc(p, &p->s64, ne));
  CREATE_TRY(dalloc(p, &p->s32, ne));
  CREATE_TRY(dalloc(p, &p->design, ne));
  CREATE_TRY(dalloc(p, &p->fixcnt, nd));
  CREATE_TRY(dalloc(p, &p->F, nd));
  CREATE_TRY(dalloc(p, &p->Fb, nd));
  CREATE_TRY(dalloc(p, &p->diag, nd));
  CREATE_TRY(dalloc(p, &p->dinv, nd));
  CREATE_TRY(dalloc(p, &p->dinv32, nd));
  CREATE_TRY(dalloc(p, &p->u, nd));
  CREATE_TRY(dalloc(p, &p->r, nd));
  CREATE_TRY(dalloc(p, &p->z, nd));
  CREATE_TRY(dalloc(p, &p->p[0], nd));
  CREATE_TRY(dalloc(p, &p->p[1], nd));
  CREATE_TRY(dalloc(p, &p->q, nd));
  if (cfg->warm_start == 2 && cfg->precond == AMT_PRECOND_MULTIGRID) {
    CREATE_TRY(dalloc(p, &p->u2, nd));
    if (hipMemset(p->u2, 0, sizeof(double) * nd) != hipSuccess)
      return bail(fail(p, AMT_EHIP, "hipMemset failed"));
  }
  if (cfg->precision == AMT_PREC_MIXED_FP32 || cfg->precision == AMT_PREC_MIXED_BF16) {
    CREATE_TRY(dalloc(p, &p->r32, nd));
    CREATE_TRY(dalloc(p, &p->z32, nd));
    CREATE_TRY(dalloc(p, &p->p32[0], nd));
    CREATE_TRY(dalloc(p, &p->p32[1], nd));
    CREATE_TRY(dalloc(p, &p->q32, nd));
    CREATE_TRY(dalloc(p, &p->d32, nd));
  }
  CREATE_TRY(dalloc(p, &p->dc, ne));
  CREATE_TRY(dalloc(p, &p->dcf, ne));
  p->tg = make_tile_grid(p->g);
  p->nb_tile = 8u * (unsigned)p->tg.per_xcd;
  p->stride_apply = (int)std::max<int64_t>({(int64_t)nblk(p->n_n), (int64_t)nblk(ne),
                                            (int64_t)p->nb_tile});
  // every operator candidate's partials fit the R_APPLY region
  for (int c = 1; c < kOpCandidates; ++c)
    p->stride_apply = std::max<int>(p->stride_apply, (int)make_op_sel(p, c).nb);
  {
    const int forced = cfg->operator_tile;
    p->op_apply = p->op_upd = p->op_post = make_op_sel(p, forced < 0 ? 0 : forced);
    p->ops_tuned = forced >= 0;
  }
  p->stride_red = std::max<int>(kMaxReduceBlocks, p->stride_apply);
  {
    // R_APPLY: 1 value per operator block; R_UPD: 2; R_INIT: 3; R_OC: kOcCand; R_MISC: 2
    const int64_t sizes[R_COUNT] = {p->stride_apply, 2LL * p->stride_red, 3LL * p->stride_red,
                                    (int64_t)kOcCand * p->stride_red, 2LL * p->stride_red};
    int64_t off = 0;
    for (int r = 0; r < R_COUNT; ++r) {
      p->region_off[r] = off;
      off += sizes[r];
    }
    CREATE_TRY(dalloc(p, &p->partials, off));
  }
  CREATE_TRY(dalloc(p, &p->scal, S_COUNT));
  CREATE_TRY(dalloc(p, &p->pcg, 2));
  CREATE_TRY(dalloc(p, &p->fcg, 1));  // flexible CG state (multigrid and learned preconditioners)
  CREATE_TRY(dalloc(p, &p->pcg_flag, 1));
  CREATE_TRY(dalloc(p, &p->oc, 1));
  CREATE_TRY(dalloc(p, &p->iter_scal, 1));

  std::vector<StencilEntry> st;
  std::vector<double> sw;
  build_stencil(cfg->filter_radius, st, sw);
  p->n_stencil = (int)st.size();
  for (const auto& o : st)
    p->stencil_reach = std::max({p->stencil_reach, std::abs(o.dx), std::abs(o.dy), std::abs(o.dz)});
  CREATE_TRY(dalloc(p, &p->stencil, st.size()));
  CREATE_TRY(dalloc(p, &p->stencil_w, sw.size()));
  CREATE_TRY(dalloc(p, &p->dbg_idx, 1));
  CREATE_TRY(dalloc(p, &p->dbg_w, 1));
  CREATE_TRY(dalloc(p, &p->dbg_cnt, 1));

  double Ke[576];
  float Kf[576];
  hex8_stiffness(cfg->E, cfg->nu, cfg->element_size, Ke);
  for (int i = 0; i < 576; ++i) Kf[i] = (float)Ke[i];
  if (!hadamard_blocks<double>(Ke, p->had64) || !hadamard_blocks<float>(Ke, p->had32))
    return bail(fail(p, AMT_EINVAL, "element stiffness lacks the cube's symmetry"));
  CREATE_TRY(copy_in(p, p->Ke64, Ke, sizeof(Ke)));
  CREATE_TRY(copy_in(p, p->Ke32, Kf, sizeof(Kf)));
  {
    // v = hi + lo, hi = bf16(v), lo = bf16(v - hi), round to nearest even from FP32
    auto bf = [](float f) {
      uint32_t u;
      std::memcpy(&u, &f, 4);
      return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
    };
    auto fb = [](uint16_t b) {
      const uint32_t u = (uint32_t)b << 16;
      float f;
      std::memcpy(&f, &u, 4);
      return f;
    };
    uint16_t kh[576], kl[576];
    for (int i = 0; i < 576; ++i) {
      kh[i] = bf(Kf[i]);
      kl[i] = bf(Kf[i] - fb(kh[i]));
    }
    CREATE_TRY(copy_in(p, p->Kh16, kh, sizeof(kh)));
    CREATE_TRY(copy_in(p, p->Kl16, kl, sizeof(kl)));
  }
  if (cfg->precond == AMT_PRECOND_MULTIGRID) CREATE_TRY(mg_build_hierarchy(p, Ke));
  if (!st.empty()) {
    CREATE_TRY(copy_in(p, p->stencil, st.data(), st.size() * sizeof(StencilEntry)));
    CREATE_TRY(copy_in(p, p->stencil_w, sw.data(), sw.size() * sizeof(double)));
  }
  if (hipMemsetAsync(p->fixcnt, 0, nd, p->stream) != hipSuccess ||
      hipMemsetAsync(p->scal, 0, sizeof(double) * S_COUNT, p->stream) != hipSuccess)
    return bail(fail(p, AMT_EHIP, "hipMemset failed"));
  // initial density: uniform volume fraction (topology_optimizer.cpp:403-405)
  hipLaunchKernelGGL(k_fill<double>, dim3(nblk(ne)), dim3(kBlock), 0, p->stream, p->rho,
                     cfg->volume_fraction, ne);
  if (hipStreamSynchronize(p->stream) != hipSuccess)
    return bail(fail(p, AMT_EHIP, "initialisation failed"));
#undef CREATE_TRY
  *out = p;
  return AMT_OK;
}

int amt_destroy(amt_problem* p) {
  if (!p) return AMT_OK;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  for (auto e : p->ev_pool) (void)hipEventDestroy(e);
  if (p->fcg_exec) (void)hipGraphExecDestroy(p->fcg_exec);
  if (p->fcg_graph) (void)hipGraphDestroy(p->fcg_graph);
  if (p->post_exec) (void)hipGraphExecDestroy(p->post_exec);
  if (p->post_graph) (void)hipGraphDestroy(p->post_graph);

  for (void* a : p->allocs) (void)hipFree(a);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return AMT_OK;
}

int amt_set_loads(amt_problem* p, const double* F, int64_t n_dof, const int64_t* fixed,
                  int64_t n_fixed) {
  if (!p) return AMT_EINVAL;
  if (!F || n_dof != p->n_dof)
    return fail(p, AMT_EINVAL,
                "force vector size mismatch: expected " + std::to_string(p->n_dof) + ", got " +
                    std::to_string(n_dof));
  if (n_fixed < 0 || (n_fixed > 0 && !fixed)) return fail(p, AMT_EINVAL, "bad fixed list");
  std::vector<uint8_t> cnt(p->n_dof, 0);
  for (int64_t t = 0; t < n_fixed; ++t) {
    const int64_t i = fixed[t];
    if (i < 0 || i >= p->n_dof)
      return fail(p, AMT_ERANGE, "Fixed DOF index out of range: " + std::to_string(i));
    if (cnt[i] == 255) return fail(p, AMT_EINVAL, "DOF listed more than 255 times");
    cnt[i]++;
  }
  for (int64_t i = 0; i < n_dof; ++i)
    if (!std::isfinite(F[i])) return fail(p, AMT_ENAN, "non-finite force");
  AMT_CHECK(copy_in(p, p->F, F, sizeof(double) * n_dof));
  AMT_CHECK(copy_in(p, p->fixcnt, cnt.data(), n_dof));
  p->loads_set = true;
  p->coeffs_ready = false;
  p->have_u = false;
  return AMT_OK;
}

int amt_set_density(amt_problem* p, const double* rho, int64_t n_e) {
  if (!p) return AMT_EINVAL;
  if (!rho || n_e != p->n_e)
    return fail(p, AMT_EINVAL,
                "Density vector size mismatch: expected " + std::to_string(p->n_e) + ", got " +
                    std::to_string(n_e));
  for (int64_t i = 0; i < n_e; ++i)
    if (!std::isfinite(rho[i])) return fail(p, AMT_ENAN, "non-finite density");
  AMT_CHECK(copy_in(p, p->rho, rho, sizeof(double) * n_e));
  p->coeffs_ready = false;
  return AMT_OK;
}

int amt_get_density(amt_problem* p, double* rho, int64_t n_e) {
  if (!p) return AMT_EINVAL;
  if (!rho || n_e != p->n_e) return fail(p, AMT_EINVAL, "density buffer size mismatch");
  return copy_out(p, rho, p->rho, sizeof(double) * n_e);
}

int amt_get_displacement(amt_problem* p, double* u, int64_t n_dof) {
  if (!p) return AMT_EINVAL;
  if (!u || n_dof != p->n_dof) return fail(p, AMT_EINVAL, "displacement buffer size mismatch");
  return copy_out(p, u, p->u, sizeof(double) * n_dof);
}

int amt_begin_run(amt_problem* p) {
  if (!p) return AMT_EINVAL;
  return build_design(p);
}

int amt_run_iteration(amt_problem* p, amt_iteration_result* out) {
  if (!p) return AMT_EINVAL;
  return run_iteration(p, out);
}

int amt_run(amt_problem* p, int max_it, amt_iteration_result* hist, int* iterations,
            int* converged, double* elapsed_s, amt_iter_cb cb, void* user) {
  if (!p) return AMT_EINVAL;
  if (max_it < 0 || (max_it > 0 && !hist)) return fail(p, AMT_EINVAL, "bad history buffer");
  const double t0 = now_s();
  AMT_CHECK(build_design(p));
  int done = 0, conv = 0;
  const int cap = std::min(max_it, p->cfg.max_iterations);
  std::vector<double> host_rho;
  for (int it = 0; it < cap; ++it) {
    amt_iteration_result R;
    AMT_CHECK(run_iteration(p, &R));
    hist[it] = R;
    done = it + 1;
    if (p->cfg.verbose)
      std::printf("%6d %14.4e %10.3f %12.2e %10d\n", it, R.compliance, R.volume_fraction,
                  R.max_density_change, R.solver_iterations);
    if (cb) {
      host_rho.resize(p->n_e);
      AMT_CHECK(copy_out(p, host_rho.data(), p->rho, sizeof(double) * p->n_e));
      if (!cb(it, host_rho.data(), p->n_e, &R, user)) break;
    }
    if (R.max_density_change < p->cfg.convergence_tol) {
      conv = 1;
      break;
    }
  }
  if (iterations) *iterations = done;
  if (converged) *converged = conv;
  if (elapsed_s) *elapsed_s = now_s() - t0;
  return AMT_OK;
}

int amt_apply_operator(amt_problem* p, const double* x, double* y, int with_bc) {
  if (!p || !x || !y) return AMT_EINVAL;
  if (with_bc && !p->loads_set) return fail(p, AMT_EINVAL, "loads not set (amt_set_loads)");
  if (!p->coeffs_ready) AMT_CHECK(prepare_coeffs(p));
  AMT_CHECK(copy_in(p, p->z, x, sizeof(double) * p->n_dof));
  launch_apply64(p, p->op_apply, p->z, p->q, with_bc ? (const double*)(p->scal + S_P) : nullptr);
  AMT_CHECK(check_launch(p, "k_apply"));
  return copy_out(p, y, p->q, sizeof(double) * p->n_dof);
}

int amt_apply_operator_prec(amt_problem* p, const double* x, double* y, int with_bc,
                            int precision) {
  if (!p || !x || !y) return AMT_EINVAL;
  if (precision == AMT_PREC_FP64) return amt_apply_operator(p, x, y, with_bc);
  if (precision != AMT_PREC_MIXED_FP32 && precision != AMT_PREC_MIXED_BF16)
    return fail(p, AMT_EINVAL, "unknown precision");
  if (with_bc && !p->loads_set) return fail(p, AMT_EINVAL, "loads not set (amt_set_loads)");
  if (!p->coeffs_ready) AMT_CHECK(prepare_coeffs(p));
  float *xf = nullptr, *yf = nullptr;
  HIP_TRY(p, hipMalloc(&xf, sizeof(float) * p->n_dof));
  if (hipMalloc(&yf, sizeof(float) * p->n_dof) != hipSuccess) {
    (void)hipFree(xf);
    return fail(p, AMT_EHIP, "hipMalloc failed");
  }
  int rc = copy_in(p, p->z, x, sizeof(double) * p->n_dof);
  if (rc == AMT_OK) {
    hipLaunchKernelGGL(k_d2f, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream, p->z, xf,
                       p->n_dof);
    const double* P = with_bc ? (const double*)(p->scal + S_P) : nullptr;
    if (precision == AMT_PREC_MIXED_BF16)
      hipLaunchKernelGGL(k_apply_mfma, dim3(p->nb_tile), dim3(kBlock), 0, p->stream, xf, yf,
                         p->s32, p->Kh16, p->Kl16, p->fixcnt, P, p->g, p->tg);
    else
      launch_apply32(p, p->op_post, xf, yf, P);
    hipLaunchKernelGGL(k_f2d, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream, yf, p->q,
                       p->n_dof);
    rc = check_launch(p, "operator");
  }
  if (rc == AMT_OK) rc = copy_out(p, y, p->q, sizeof(double) * p->n_dof);
  (void)hipFree(xf);
  (void)hipFree(yf);
  return rc;
}

int amt_get_diagonal(amt_problem* p, double* diag, int with_bc, double* penalty) {
  if (!p || !diag) return AMT_EINVAL;
  if (!p->coeffs_ready) AMT_CHECK(prepare_coeffs(p));
  std::vector<double> dinv(p->n_dof);
  if (with_bc) {
    AMT_CHECK(copy_out(p, dinv.data(), p->dinv, sizeof(double) * p->n_dof));
    for (int64_t i = 0; i < p->n_dof; ++i) diag[i] = 1.0 / dinv[i];
  } else {
    AMT_CHECK(copy_out(p, diag, p->diag, sizeof(double) * p->n_dof));
  }
  if (penalty) AMT_CHECK(copy_out(p, penalty, p->scal + S_P, sizeof(double)));
  return AMT_OK;
}

int amt_solve(amt_problem* p, double* u, amt_solve_stats* stats) {
  if (!p) return AMT_EINVAL;
  amt_solve_stats st;
  std::memset(&st, 0, sizeof(st));
  const double t0 = now_s();
  AMT_CHECK(do_solve(p, &st));
  HIP_TRY(p, hipStreamSynchronize(p->stream));
  st.solve_time_s = now_s() - t0;
  if (stats) *stats = st;
  if (u) return copy_out(p, u, p->u, sizeof(double) * p->n_dof);
  return AMT_OK;
}

int amt_apply_multigrid(amt_problem* p, const double* r, double* z) {
  if (!p || !r || !z) return AMT_EINVAL;
  if (p->cfg.precond != AMT_PRECOND_MULTIGRID || p->mg.empty())
    return fail(p, AMT_EINVAL, "not a multigrid handle");
  if (!p->loads_set) return fail(p, AMT_EINVAL, "loads not set (amt_set_loads)");
  if (!p->coeffs_ready) AMT_CHECK(prepare_coeffs(p));
  AMT_CHECK(tune_ops(p));
  // the level-0 sweeps return at once once a solve is done: clear the flag first
  HIP_TRY(p, hipMemsetAsync((char*)p->fcg + offsetof(FcgState, done), 0, sizeof(int), p->stream));
  AMT_CHECK(copy_in(p, p->r, r, sizeof(double) * p->n_dof));
  mg_vcycle(p, p->r, p->z, nullptr);
  AMT_CHECK(check_launch(p, "multigrid V-cycle"));
  return copy_out(p, z, p->z, sizeof(double) * p->n_dof);
}

int amt_sensitivities(amt_problem* p, const double* u, double* dc) {
  if (!p || !u || !dc) return AMT_EINVAL;
  AMT_CHECK(copy_in(p, p->u, u, sizeof(double) * p->n_dof));
  p->have_u = true;
  hipLaunchKernelGGL(k_sens, dim3(nblk(p->n_e)), dim3(kBlock), 0, p->stream, p->u, p->rho,
                     p->Ke64, p->g, p->cfg.penalty, p->dc);
  AMT_CHECK(check_launch(p, "k_sens"));
  return copy_out(p, dc, p->dc, sizeof(double) * p->n_e);
}

int amt_filter(amt_problem* p, const double* dc, double* dcf) {
  if (!p || !dc || !dcf) return AMT_EINVAL;
  AMT_CHECK(copy_in(p, p->dc, dc, sizeof(double) * p->n_e));
  launch_filter(p);
  AMT_CHECK(check_launch(p, "filter"));
  return copy_out(p, dcf, p->dcf, sizeof(double) * p->n_e);
}

int amt_filter_neighbors(amt_problem* p, int64_t e, int64_t* idx, double* w, int cap,
                         int* count) {
  if (!p || !count || cap < 0 || (cap > 0 && (!idx || !w))) return AMT_EINVAL;
  if (e < 0 || e >= p->n_e) return fail(p, AMT_ERANGE, "element index out of range");
  if (cap > 0) {
    // scratch large enough for this request
    (void)hipFree(p->dbg_idx);
    (void)hipFree(p->dbg_w);
    p->allocs.erase(std::remove(p->allocs.begin(), p->allocs.end(), (void*)p->dbg_idx),
                    p->allocs.end());
    p->allocs.erase(std::remove(p->allocs.begin(), p->allocs.end(), (void*)p->dbg_w),
                    p->allocs.end());
    p->dbg_idx = nullptr;
    p->dbg_w = nullptr;
    AMT_CHECK(dalloc(p, &p->dbg_idx, cap));
    AMT_CHECK(dalloc(p, &p->dbg_w, cap));
  }
  hipLaunchKernelGGL(k_filter_list, dim3(1), dim3(64), 0, p->stream, e, p->stencil,
                     p->stencil_w, p->n_stencil, p->g, p->dbg_idx, p->dbg_w, cap, p->dbg_cnt);
  AMT_CHECK(check_launch(p, "k_filter_list"));
  AMT_CHECK(copy_out(p, count, p->dbg_cnt, sizeof(int)));
  const int m = std::min(cap, *count);
  if (m > 0) {
    AMT_CHECK(copy_out(p, idx, p->dbg_idx, sizeof(int64_t) * m));
    AMT_CHECK(copy_out(p, w, p->dbg_w, sizeof(double) * m));
  }
  return AMT_OK;
}

int amt_oc_update(amt_problem* p, const double* dcf, double* max_change, int* passes) {
  if (!p || !dcf) return AMT_EINVAL;
  AMT_CHECK(copy_in(p, p->dcf, dcf, sizeof(double) * p->n_e));
  OcState hs;
  AMT_CHECK(oc_update(p, &hs));
  if (max_change) *max_change = hs.change;
  if (passes) *passes = hs.passes;
  return AMT_OK;
}

int amt_unet_create_for(amt_problem* p, uint64_t seed, amt_unet** out) {
  if (!p || !out) return AMT_EINVAL;
  int px, py, pz;
  nn_dims(p, px, py, pz);
  return amt_unet_create(px, py, pz, seed, out);
}

int amt_attach_unet(amt_problem* p, amt_unet* net) {
  if (!p) return AMT_EINVAL;
  if (net) AMT_CHECK(nn_check(p, net));
  p->net = net;
  if (net) {
    AMT_CHECK(nn_alloc(p));
    net->stream = p->stream;
    p->coeffs_ready = false;
  }
  return AMT_OK;
}

int amt_unet_collect(amt_problem* p, int max_samples, int* n_total) {
  if (!p || max_samples < 0) return AMT_EINVAL;
  if (!p->loads_set) return fail(p, AMT_EINVAL, "loads not set (amt_set_loads)");
  if (!p->coeffs_ready) AMT_CHECK(prepare_coeffs(p));
  int taken = 0;
  auto snap = [&](const double* r) -> int {
    AMT_CHECK(bank_push(p, r));
    ++taken;
    return AMT_OK;
  };
  if (max_samples > 0) {
    // plain FP64 Jacobi-PCG from u = 0; the residual after every 32 passes is a sample
    hipLaunchKernelGGL(k_fill<double>, dim3(nblk(p->n_dof)), dim3(kBlock), 0, p->stream, p->u,
                       0.0, p->n_dof);
    HIP_TRY(p, hipMemcpyAsync(p->r, p->Fb, sizeof(double) * p->n_dof, hipMemcpyDeviceToDevice,
                              p->stream));
    AMT_CHECK(snap(p->r));
    int rc_snap = AMT_OK;
    std::function<bool()> hook = [&]() {
      if (taken >= max_samples) return false;
      rc_snap = snap(p->r);
      return rc_snap == AMT_OK && taken < max_samples;
    };
    PcgState hs;
    AMT_CHECK((pcg_run<double, double, double>(p, p->Fb, p->u, p->r, p->z, p->p, p->q, p->s64,
                                               p->Ke64, p->dinv, p->cfg.pcg_tol,
                                               p->cfg.pcg_maxit, &hs, &hook)));
    AMT_CHECK(rc_snap);
    p->have_u = false;
  }
  HIP_TRY(p, hipStreamSynchronize(p->stream));
  if (n_total) *n_total = (int)p->bank.size();
  return AMT_OK;
}

int amt_unet_collect_warm(amt_problem* p, int max_samples, int* n_total) {
  if (!p || max_samples < 0) return AMT_EINVAL;
  if (p->cfg.precond != AMT_PRECOND_MULTIGRID)
    return fail(p, AMT_EINVAL, "warm-start samples come from multigrid solves");
  p->warm_collect = max_samples;
  if (n_total) *n_total = (int)p->bank.size();
  return AMT_OK;
}

int amt_unet_train(amt_problem* p, amt_unet* net, int steps, double lr, double* losses,
                   double* jacobi_losses) {
  if (!p || !net || steps < 0) return AMT_EINVAL;
  AMT_CHECK(nn_check(p, net));
  if (p->bank.empty()) return fail(p, AMT_EINVAL, "no training samples (amt_unet_collect)");
  const double* last_rho = nullptr;
  AMT_CHECK(nn_alloc(p));
  amt_unet* saved = p->net;
  p->net = net;
  if (!p->coeffs_ready) AMT_CHECK(prepare_coeffs(p));
  const bool want = losses || jacobi_losses;
  double* lt = nullptr;
  if (want) AMT_CHECK(dalloc(p, &lt, 5 * (size_t)std::max(steps, 1)));
  for (int s = 0; s < steps; ++s) {
    const auto& smp = p->bank[s % p->bank.size()];
    if (smp.rho != last_rho) {  // the sample's operator, Jacobi scaling and density feature
      sample_coeffs(p, smp.rho);
      last_rho = smp.rho;
    }
    nn_train_step(p, net, smp.r, (float)lr, lt ? lt + 5 * s : nullptr);
  }
  p->coeffs_ready = false;  // the density is the last sample's; the next solve rebuilds
  AMT_CHECK(check_launch(p, "unet train"));
  if (want) {
    std::vector<double> h(5 * (size_t)steps);
    AMT_CHECK(copy_out(p, h.data(), lt, sizeof(double) * h.size()));
    for (int s = 0; s < steps; ++s) {
      const double* t = &h[5 * (size_t)s];
      if (losses) losses[s] = (t[0] - 2.0 * t[1]) / t[2];
      if (jacobi_losses) jacobi_losses[s] = (t[3] - 2.0 * t[4]) / t[2];
    }
  }
  HIP_TRY(p, hipStreamSynchronize(p->stream));
  p->net = saved;
  if (saved) saved->stream = p->stream;
  return AMT_OK;
}

int amt_sync(amt_problem* p) {
  if (!p) return AMT_EINVAL;
  HIP_TRY(p, hipStreamSynchronize(p->stream));
  return AMT_OK;
}

int amt_device_info(char* name, int cap, int* n_cu) {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess)
    return fail(nullptr, AMT_EHIP, "no HIP device");
  if (name && cap > 0) {
    std::snprintf(name, cap, "%s (%s)", prop.name, prop.gcnArchName);
  }
  if (n_cu) *n_cu = prop.multiProcessorCount;
  return AMT_OK;
}

int amt_operator_info(amt_problem* p, int kind, int* candidate, int* tj, int* tk, int* planes,
                      double* us) {
  if (!p || kind < 0 || kind > 2) return AMT_EINVAL;
  const amt_problem::OpSel& s = kind == 0 ? p->op_apply : kind == 1 ? p->op_upd : p->op_post;
  if (candidate) *candidate = s.cand;
  if (tj) *tj = s.shape < 0 ? kTJ : kStreamShape[s.shape][0];
  if (tk) *tk = s.shape < 0 ? kTK : kStreamShape[s.shape][1];
  if (planes) *planes = s.shape < 0 ? kTI : s.sg.L;
  if (us) *us = s.us;
  return AMT_OK;
}

int amt_operator_timings(amt_problem* p, int kind, double* us, int n) {
  if (!p || kind < 0 || kind > 2 || n < 0 || (n > 0 && !us)) return AMT_EINVAL;
  for (int c = 0; c < std::min(n, kOpCandidates); ++c) us[c] = p->tune_us[kind][c];
  return AMT_OK;
}

int amt_profile_enable(amt_problem* p, int on) {
  if (!p) return AMT_EINVAL;
  p->prof_on = on != 0;
  p->prof_ms = p->prof_ms_upd = 0.0;
  p->prof_launches = p->prof_launches_upd = 0;
  p->prof_tick = 0;
  if (p->prof_on && p->ev_pool.empty()) {
    p->ev_pool.resize(2 * 4096);
    p->ev_kind.assign(4096, 0);
    for (auto& e : p->ev_pool) HIP_TRY(p, hipEventCreate(&e));
  }
  return AMT_OK;
}

int amt_profile_read(amt_problem* p, int64_t* launches, double* total_ms) {
  return amt_profile_read_kind(p, 0, launches, total_ms);
}

int amt_profile_read_kind(amt_problem* p, int kind, int64_t* launches, double* total_ms) {
  if (!p) return AMT_EINVAL;
  if (kind < 0 || kind > 1) return fail(p, AMT_ERANGE, "profile kind must be 0 or 1");
  prof_harvest(p);
  if (launches) *launches = kind ? p->prof_launches_upd : p->prof_launches;
  if (total_ms) *total_ms = kind ? p->prof_ms_upd : p->prof_ms;
  return AMT_OK;
}

}  // extern "C"

// ---- OC update on caller buffers, without allocation ----------------------------------------
// The kernels' scratch (sweep and bound partials, final partials, bisection state) lives in one
// file-scope device array. Its host-side name is only a stand-in for the symbol: kernels get the
// device address that hipGetSymbolAddress returns, looked up once per device and kept.
namespace {
constexpr int kNaStride = kOcBlocks;                           // partials per value
constexpr size_t kNaSweep = (size_t)kOcCand * kNaStride;       // candidate volumes (>= 4
                                                               // bound values)
constexpr size_t kNaFinal = 2 * (size_t)kNaStride;             // max |drho|, volume (count)
constexpr size_t kNaState = kNaSweep + kNaFinal;               // OcState offset, in doubles
constexpr size_t kNaDoubles = kNaState + (sizeof(OcState) + 7) / 8;
static_assert(kOcCand >= 4, "the bound partials share the sweep region");
constexpr int kNaMaxDevices = 64;
std::mutex g_na_mutex;
double* g_na_ptr[kNaMaxDevices] = {nullptr};
}  // namespace

static __device__ __align__(256) double scratch[kNaDoubles];

extern "C" int amt_oc_update_noalloc(double* rho, const double* dcf, const uint8_t* design,
                                     int64_t n_e, double volume_fraction, double move_limit,
                                     double rho_min, double* result, void* stream) {
  if (!rho || !dcf || !design || !result)
    return set_global_error(AMT_EINVAL, "amt_oc_update_noalloc: null buffer");
  if (n_e <= 0 || n_e > AMT_OC_NOALLOC_MAX_ELEMENTS)
    return set_global_error(AMT_EINVAL,
                            "amt_oc_update_noalloc: n_e = " + std::to_string(n_e) +
                                " outside [1, 2^31] (the limit of the static scratch)");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNaMaxDevices)
    return set_global_error(AMT_EHIP, "amt_oc_update_noalloc: no usable HIP device");
  double* S = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_na_mutex);
    if (!g_na_ptr[dev]) {
      void* ptr = nullptr;
      const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !ptr)
        return set_global_error(AMT_EHIP, std::string("amt_oc_update_noalloc: "
                                                      "hipGetSymbolAddress: ") +
                                              hipGetErrorString(e));
      g_na_ptr[dev] = (double*)ptr;
    }
    S = g_na_ptr[dev];
  }
  hipStream_t s = (hipStream_t)stream;
  double* part = S;                      // bound partials, then the sweep's
  double* fin = S + kNaSweep;            // design count, then the final partials
  OcState* st = (OcState*)(S + kNaState);
  const unsigned nb = std::min<unsigned>(nred(n_e), kOcBlocks);
  hipLaunchKernelGGL(k_design_count, dim3(nb), dim3(kBlock), 0, s, design, n_e, fin, kNaStride);
  hipLaunchKernelGGL(k_oc_state_init, dim3(1), dim3(kBlock), 0, s, st, (const double*)fin,
                     (int)nb, kNaStride, n_e, volume_fraction);
  hipLaunchKernelGGL(k_oc_bounds, dim3(nb), dim3(kBlock), 0, s, rho, dcf, design, n_e,
                     move_limit, rho_min, part, kNaStride);
  hipLaunchKernelGGL(k_oc_bounds_final, dim3(1), dim3(kBlock), 0, s, st, (const double*)part,
                     (int)nb, kNaStride);
  for (int sw = 0; sw < (100 + kOcLevels - 1) / kOcLevels + 1; ++sw) {
    hipLaunchKernelGGL((k_oc_sweep<kOcLevels>), dim3(nb), dim3(kBlock), 0, s, rho, dcf, design,
                       n_e, move_limit, rho_min, st, part, kNaStride);
    hipLaunchKernelGGL((k_oc_walk<kOcLevels>), dim3(1), dim3(kBlock), 0, s, st,
                       (const double*)part, (int)nb, kNaStride);
  }
  hipLaunchKernelGGL(k_oc_final, dim3(nb), dim3(kBlock), 0, s, rho, dcf, design, n_e,
                     move_limit, rho_min, st, fin, kNaStride);
  hipLaunchKernelGGL(k_oc_final_reduce, dim3(1), dim3(kBlock), 0, s, st, (const double*)fin,
                     (int)nb, kNaStride);
  hipLaunchKernelGGL(k_oc_state_out, dim3(1), dim3(64), 0, s, (const OcState*)st, result);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_global_error(AMT_EHIP, std::string("amt_oc_update_noalloc: launch: ") +
                                          hipGetErrorString(e));
  return AMT_OK;
}
