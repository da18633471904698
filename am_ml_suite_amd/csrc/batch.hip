// batch.hip — the batched SIMP engine (dataset generation of many uneven problems) and its
// C-ABI (include/amt.h, amt_batch_*). No counterpart in the reference, which optimises one
// problem per TopologyOptimizer (topology_optimizer.cpp:391-520); each problem here follows
// those semantics (consistent density index, Jacobi-PCG with penalty BCs, filter, OC).
//
// All problems advance in lockstep, one OC iteration at a time. Every phase is one persistent
// launch over the items of all problems (batch.hpp). The pre-solve phase (SIMP scales,
// diagonal, penalty, Jacobi, F_bc), each PCG (re)start, a chunk of kBChunk PCG iterations and
// the post-solve phase (compliance, sensitivities, filter, all OC bisection sweeps and the
// density update) are captured once as hipGraphs and replayed; the host only polls the
// per-problem solver flags between chunk replays. Problems that converge drop out of the
// chunk at once (their items exit at the top). PCG runs its matvec in T (FP32 in the mixed
// mode) with FP64 u, r and reductions. After every chunk the FP64 residual F_bc - K u replaces
// the recursive one (van der Vorst-Ye residual replacement: the search direction is kept), so
// convergence is judged on the true residual and the FP32 matvec error cannot let the two
// drift apart.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "../../include/amt.h"
#include "batch.hpp"
#include "host_util.hpp"

namespace amt {
void hex8_stiffness(double E, double nu, double h, double* Ke);
int set_global_error(int code, const std::string& msg);
}  // namespace amt

using namespace amt;

namespace {

constexpr int kBChunk = 16;        // PCG iterations per chunk graph (even: parity returns)
constexpr int kBMgWorkers = 8;     // host threads of the multigrid batch
constexpr int kBPersist = 2048;    // blocks of a persistent launch (8 per CU)
constexpr int kBOcLevels = 3;  // bisection levels per OC sweep (7 candidates)
constexpr int kBOcCand = (1 << kBOcLevels) - 1;
// Matvec candidates of the lockstep solver: 0 the LDS gather tile (BK_TILE items); 1 + 3 s + l
// the element-column streaming tile of shape kBShape[s] (threads along j x k) with kBPlanes[l]
// node planes per tile. Chosen per batch by timing (amt_config.operator_tile = -1) or forced.
constexpr int kBShapes = 3, kBPlaneCounts = 3, kBCands = 1 + kBShapes * kBPlaneCounts;
constexpr int kBShapeTJ[kBShapes] = {16, 8, 4}, kBShapeTK[kBShapes] = {16, 32, 64};
constexpr int kBPlanes[kBPlaneCounts] = {2, 4, 8};

int bfail(int code, const std::string& msg) { return amt::set_global_error(code, msg); }

#define B_HIP(call)                                                                    \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return bfail(AMT_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));       \
  } while (0)
#define B_CHECK(call)              \
  do {                             \
    int rc_ = (call);              \
    if (rc_ != AMT_OK) return rc_; \
  } while (0)

}  // namespace

struct amt_batch {
  int np = 0;
  // multigrid mode: one single-problem handle (and HIP stream) per problem, driven by
  // kBMgWorkers host threads, so the small problems' solver launches overlap on the device
  std::vector<amt_problem*> mg;
  std::vector<uint8_t> mg_begun;
  bool mixed = true, warm = true;
  std::vector<amt_config> cfg;
  std::vector<BProb> hp;
  BProb* D = nullptr;
  int n_items[BK_KINDS] = {0, 0, 0, 0};
  int* ip[BK_KINDS] = {nullptr, nullptr, nullptr, nullptr};
  int64_t n_e = 0, n_n = 0, n_dof = 0;
  // packed problem data
  double *rho = nullptr, *s64 = nullptr, *dc = nullptr, *dcf = nullptr;
  float* s32 = nullptr;
  uint8_t *design = nullptr, *fixcnt = nullptr;
  double *u = nullptr, *r = nullptr, *F = nullptr, *Fb = nullptr, *dinv = nullptr,
         *diag = nullptr;
  void *z = nullptr, *p0 = nullptr, *p1 = nullptr, *q = nullptr;  // T
  double* Ke64 = nullptr;
  float* Ke32 = nullptr;
  StencilEntry* st = nullptr;
  double* sw = nullptr;
  // per problem
  BState *S = nullptr, *flags = nullptr;
  OcState* oc = nullptr;
  double *maxdiag = nullptr, *bb = nullptr, *pen = nullptr, *ndesign = nullptr,
         *compl_ = nullptr;
  // partials
  double *part_tile = nullptr, *part_dof = nullptr, *part_elem = nullptr, *part_node = nullptr;
  // matvec candidates: the chosen one (-1 until the first run tunes it), the tuner's times,
  // per candidate the items (count, owners, per-problem stream grids) and per problem the
  // range of its partials (item0, nitem); the Hadamard blocks per Ke0 table
  int op_cand = -1;
  double op_us[kBCands] = {};
  int nS[kBCands] = {};
  int* ipS[kBCands] = {};
  BStream* DS[kBCands] = {};
  int* ar[kBCands] = {};
  bool had_ok = false;
  HadK<double>* HK64 = nullptr;
  HadK<float>* HK32 = nullptr;
  double* part_apply = nullptr;  // first-half partials p.q, the largest candidate's item count
  hipStream_t stream = nullptr;
  hipGraphExec_t g_pre = nullptr, g_start0 = nullptr, g_start1 = nullptr, g_chunk = nullptr,
                 g_post = nullptr;
  std::vector<uint8_t> loads_set;
  bool design_ready = false;
  std::vector<double> ndesign_h;
  std::vector<void*> allocs;
};

namespace {

template <typename T>
int balloc(amt_batch* b, T** p, size_t n) {
  void* v = nullptr;
  if (hipMalloc(&v, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
    return bfail(AMT_EHIP, "hipMalloc failed (batch buffers)");
  b->allocs.push_back(v);
  *p = (T*)v;
  return AMT_OK;
}

unsigned grid_of(int n_items) {
  const int g = std::min(((n_items + 7) / 8) * 8, kBPersist);
  return (unsigned)std::max(g, 8);
}

void destroy(amt_batch* b) {
  if (!b) return;
  for (amt_problem* h : b->mg) amt_destroy(h);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (hipGraphExec_t g : {b->g_pre, b->g_start0, b->g_start1, b->g_chunk, b->g_post})
    if (g) (void)hipGraphExecDestroy(g);
  for (void* a : b->allocs) (void)hipFree(a);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

// ---- phases (enqueued on b->stream; captured into graphs) ----

template <typename T>
void enqueue_pre(amt_batch* b) {
  const int ne = b->n_items[BK_ELEM], nn = b->n_items[BK_NODE], nd = b->n_items[BK_DOF];
  hipLaunchKernelGGL(k_b_simp, dim3(grid_of(ne)), dim3(kBlock), 0, b->stream, b->D,
                     b->ip[BK_ELEM], ne, b->rho, b->s64, b->s32);
  hipLaunchKernelGGL(k_b_diag, dim3(grid_of(nn)), dim3(kBlock), 0, b->stream, b->D,
                     b->ip[BK_NODE], nn, b->s64, b->Ke64, b->diag, b->part_node);
  hipLaunchKernelGGL((k_b_reduce<1, true>), dim3(b->np), dim3(kBlock), 0, b->stream, b->D,
                     (int)BK_NODE, b->part_node, nn, b->maxdiag, 1);
  hipLaunchKernelGGL(k_b_dinv, dim3(grid_of(nd)), dim3(kBlock), 0, b->stream, b->D,
                     b->ip[BK_DOF], nd, b->maxdiag, b->diag, b->fixcnt, b->F, b->pen, b->dinv,
                     b->Fb, b->part_dof);
  hipLaunchKernelGGL((k_b_reduce<1>), dim3(b->np), dim3(kBlock), 0, b->stream, b->D,
                     (int)BK_DOF, b->part_dof, nd, b->bb, 1);
  if (!b->warm)
    hipLaunchKernelGGL(k_fill<double>, dim3((unsigned)((b->n_dof + kBlock - 1) / kBlock)),
                       dim3(kBlock), 0, b->stream, b->u, 0.0, b->n_dof);
}

// Residual replacement after a chunk loop: FP64 residual from u, per-problem decision
// (k_b_replace), last directions moved to buffer 0.
template <typename T>
void enqueue_replace(amt_batch* b) {
  const int nt = b->n_items[BK_TILE], nd = b->n_items[BK_DOF];
  hipLaunchKernelGGL((k_b_residual<T>), dim3(grid_of(nt)), dim3(kBlock), 0, b->stream, b->D,
                     b->ip[BK_TILE], nt, b->u, b->Fb, b->r, (T*)b->z, b->dinv, b->s64, b->Ke64,
                     b->fixcnt, b->pen, b->part_tile);
  hipLaunchKernelGGL(k_b_replace, dim3(b->np), dim3(kBlock), 0, b->stream, 0, b->D,
                     b->part_tile, nt, b->S, b->part_dof, nd, b->flags);
  hipLaunchKernelGGL((k_b_carry_p<T>), dim3(grid_of(nd)), dim3(kBlock), 0, b->stream, b->D,
                     b->ip[BK_DOF], nd, b->flags, (T*)b->p0, (const T*)b->p1);
}

// FP64 residual from u, PCG state per problem, p_old = 0.
template <typename T>
void enqueue_start(amt_batch* b, int first) {
  const int nt = b->n_items[BK_TILE];
  hipLaunchKernelGGL((k_b_residual<T>), dim3(grid_of(nt)), dim3(kBlock), 0, b->stream, b->D,
                     b->ip[BK_TILE], nt, b->u, b->Fb, b->r, (T*)b->z, b->dinv, b->s64, b->Ke64,
                     b->fixcnt, b->pen, b->part_tile);
  hipLaunchKernelGGL(k_b_start, dim3(b->np), dim3(kBlock), 0, b->stream, b->D, b->part_tile,
                     nt, b->bb, first, b->S, b->flags);
  hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)((b->n_dof + kBlock - 1) / kBlock)),
                     dim3(kBlock), 0, b->stream, (T*)b->p0, T(0), b->n_dof);
}

// Items of matvec candidate c.
int cand_items(const amt_batch* b, int c) { return c == 0 ? b->n_items[BK_TILE] : b->nS[c]; }

// First half of an iteration of parity par on matvec candidate c.
template <typename T>
void launch_apply(amt_batch* b, int c, int par) {
  const int nd = b->n_items[BK_DOF], n = cand_items(b, c);
  const T* s = std::is_same<T, float>::value ? (const T*)b->s32 : (const T*)b->s64;
  T* pp[2] = {(T*)b->p0, (T*)b->p1};
  if (c == 0) {
    const T* Ke = std::is_same<T, float>::value ? (const T*)b->Ke32 : (const T*)b->Ke64;
    hipLaunchKernelGGL((k_b_apply<T>), dim3(grid_of(n)), dim3(kBlock), 0, b->stream, par, b->D,
                       b->ip[BK_TILE], n, (const T*)b->z, pp[par], pp[par ^ 1], (T*)b->q, s, Ke,
                       b->fixcnt, b->pen, b->S, b->part_dof, nd, b->part_apply);
    return;
  }
  const HadK<T>* HK = std::is_same<T, float>::value ? (const HadK<T>*)b->HK32
                                                     : (const HadK<T>*)b->HK64;
#define AMT_B_APPLY_S(TJ, TK)                                                                 \
  hipLaunchKernelGGL((k_b_apply_s<T, TJ, TK>), dim3(grid_of(n)), dim3(kBlock), 0, b->stream, \
                     par, b->D, b->DS[c], b->ipS[c], n, (const T*)b->z, pp[par], pp[par ^ 1], \
                     (T*)b->q, s, HK, b->fixcnt, b->pen, b->S, b->part_dof, nd, b->part_apply)
  switch ((c - 1) / kBPlaneCounts) {
    case 0: AMT_B_APPLY_S(16, 16); break;
    case 1: AMT_B_APPLY_S(8, 32); break;
    default: AMT_B_APPLY_S(4, 64); break;
  }
#undef AMT_B_APPLY_S
}

template <typename T>
void enqueue_chunk(amt_batch* b) {
  const int nd = b->n_items[BK_DOF], c = b->op_cand;
  T* pp[2] = {(T*)b->p0, (T*)b->p1};
  for (int t = 0; t < kBChunk; ++t) {
    const int par = t & 1;
    launch_apply<T>(b, c, par);
    hipLaunchKernelGGL((k_b_update<T>), dim3(grid_of(nd)), dim3(kBlock), 0, b->stream, par ^ 1,
                       b->D, b->ip[BK_DOF], nd, b->u, b->r, (T*)b->z, pp[par ^ 1],
                       (const T*)b->q, b->dinv, b->S, b->part_apply, b->ar[c],
                       cand_items(b, c), b->part_dof);
  }
  hipLaunchKernelGGL(k_b_check, dim3(1), dim3(kBlock), 0, b->stream, 0, b->D, b->np, b->S,
                     b->part_dof, nd, b->flags);
}

void enqueue_post(amt_batch* b) {
  const int ne = b->n_items[BK_ELEM], nd = b->n_items[BK_DOF];
  // compliance with the unmodified forces (topology_optimizer.cpp:338)
  hipLaunchKernelGGL(k_b_dot, dim3(grid_of(nd)), dim3(kBlock), 0, b->stream, b->D,
                     b->ip[BK_DOF], nd, b->u, b->F, b->part_dof);
  hipLaunchKernelGGL((k_b_reduce<1>), dim3(b->np), dim3(kBlock), 0, b->stream, b->D,
                     (int)BK_DOF, b->part_dof, nd, b->compl_, 1);
  hipLaunchKernelGGL(k_b_sens, dim3(grid_of(ne)), dim3(kBlock), 0, b->stream, b->D,
                     b->ip[BK_ELEM], ne, b->u, b->rho, b->Ke64, b->dc);
  hipLaunchKernelGGL(k_b_filter, dim3(grid_of(ne)), dim3(kBlock), 0, b->stream, b->D,
                     b->ip[BK_ELEM], ne, b->dc, b->rho, b->st, b->sw, b->dcf);
  hipLaunchKernelGGL(k_b_oc_init, dim3((b->np + 63) / 64), dim3(64), 0, b->stream, b->D, b->np,
                     b->ndesign, b->oc);
  // ceil(100 / LEV) + 1 sweeps always suffice: each advances LEV passes or finishes
  for (int sw = 0; sw < (100 + kBOcLevels - 1) / kBOcLevels + 1; ++sw) {
    hipLaunchKernelGGL((k_b_oc_sweep<kBOcLevels>), dim3(grid_of(ne)), dim3(kBlock), 0,
                       b->stream, b->D, b->ip[BK_ELEM], ne, b->rho, b->dcf, b->design, b->oc,
                       b->part_elem);
    hipLaunchKernelGGL((k_b_oc_walk<kBOcLevels>), dim3(b->np), dim3(kBlock), 0, b->stream, b->D,
                       b->oc, b->part_elem, ne);
  }
  hipLaunchKernelGGL(k_b_oc_final, dim3(grid_of(ne)), dim3(kBlock), 0, b->stream, b->D,
                     b->ip[BK_ELEM], ne, b->rho, b->dcf, b->design, b->oc, b->part_elem);
  hipLaunchKernelGGL(k_b_oc_final_reduce, dim3(b->np), dim3(kBlock), 0, b->stream, b->D, b->oc,
                     b->part_elem, ne);
}

template <typename F>
int capture(amt_batch* b, hipGraphExec_t* out, F enqueue) {
  if (*out) return AMT_OK;
  hipGraph_t g = nullptr;
  B_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  enqueue();
  const hipError_t e1 = hipStreamEndCapture(b->stream, &g);
  if (e1 != hipSuccess) return bfail(AMT_EHIP, std::string("graph capture: ") + hipGetErrorString(e1));
  const hipError_t e2 = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e2 != hipSuccess) {
    *out = nullptr;
    return bfail(AMT_EHIP, std::string("graph instantiate: ") + hipGetErrorString(e2));
  }
  return AMT_OK;
}

int build_design(amt_batch* b) {
  const int ne = b->n_items[BK_ELEM];
  hipLaunchKernelGGL(k_b_design, dim3(grid_of(ne)), dim3(kBlock), 0, b->stream, b->D,
                     b->ip[BK_ELEM], ne, b->rho, b->design, b->part_elem);
  hipLaunchKernelGGL((k_b_reduce<1>), dim3(b->np), dim3(kBlock), 0, b->stream, b->D,
                     (int)BK_ELEM, b->part_elem, ne, b->ndesign, 1);
  b->ndesign_h.assign(b->np, 0.0);
  B_HIP(hipGetLastError());
  B_HIP(hipMemcpyAsync(b->ndesign_h.data(), b->ndesign, sizeof(double) * b->np,
                       hipMemcpyDeviceToHost, b->stream));
  B_HIP(hipStreamSynchronize(b->stream));
  b->design_ready = true;
  return AMT_OK;
}

int read_flags(amt_batch* b, std::vector<BState>& f, int* active) {
  B_HIP(hipMemcpyAsync(f.data(), b->flags, sizeof(BState) * b->np, hipMemcpyDeviceToHost,
                       b->stream));
  B_HIP(hipStreamSynchronize(b->stream));
  int a = 0;
  for (const BState& s : f) a += !s.done;
  *active = a;
  return AMT_OK;
}

// The matvec tuner: each candidate's first-half launch timed on this batch (one launch to warm
// up, then the mean of three), every problem in a fresh state so that every tile runs; the
// fastest is kept. The state slots it writes are rewritten by the next solve's start.
template <typename T>
int tune_apply(amt_batch* b) {
  BState fresh;
  std::memset(&fresh, 0, sizeof(fresh));
  fresh.fresh = 1;
  fresh.maxit = 1;
  std::vector<BState> st(2 * (size_t)b->np, fresh);
  B_HIP(hipMemcpyAsync(b->S, st.data(), sizeof(BState) * st.size(), hipMemcpyHostToDevice,
                       b->stream));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  B_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return bfail(AMT_EHIP, "hipEventCreate failed");
  }
  int rc = AMT_OK, best = 0;
  for (int c = 0; c < kBCands && rc == AMT_OK; ++c) {
    b->op_us[c] = 0.0;
    if (c > 0 && !b->had_ok) continue;
    launch_apply<T>(b, c, 0);
    (void)hipEventRecord(e0, b->stream);
    for (int r = 0; r < 3; ++r) launch_apply<T>(b, c, 0);
    (void)hipEventRecord(e1, b->stream);
    float ms = 0.f;
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      rc = bfail(AMT_EHIP, "matvec tuning failed");
      break;
    }
    b->op_us[c] = 1e3 * ms / 3.0;
    if (b->op_us[c] < b->op_us[best]) best = c;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc == AMT_OK) b->op_cand = best;
  return rc;
}

template <typename T>
int run_impl(amt_batch* b, int iterations, amt_iteration_result* hist, double* elapsed) {
  for (int i = 0; i < b->np; ++i)
    if (!b->loads_set[i])
      return bfail(AMT_EINVAL, "loads not set for problem " + std::to_string(i));
  if (!b->design_ready) B_CHECK(build_design(b));
  if (b->op_cand < 0) B_CHECK(tune_apply<T>(b));  // before the chunk graph is captured
  B_CHECK(capture(b, &b->g_pre, [&] { enqueue_pre<T>(b); }));
  B_CHECK(capture(b, &b->g_start0, [&] { enqueue_start<T>(b, 1); }));
  B_CHECK(capture(b, &b->g_start1, [&] { enqueue_replace<T>(b); }));
  B_CHECK(capture(b, &b->g_chunk, [&] { enqueue_chunk<T>(b); }));
  B_CHECK(capture(b, &b->g_post, [&] { enqueue_post(b); }));
  std::vector<BState> f(b->np);
  std::vector<OcState> oc(b->np);
  std::vector<double> compl_(b->np);
  const double t_begin = now_s();
  for (int it = 0; it < iterations; ++it) {
    const double t0 = now_s();
    B_HIP(hipGraphLaunch(b->g_pre, b->stream));
    B_HIP(hipStreamSynchronize(b->stream));
    const double t1 = now_s();
    // fresh start, then chunks of kBChunk iterations, each followed by residual replacement
    int chunks = 0, active = 0;
    B_HIP(hipGraphLaunch(b->g_start0, b->stream));
    B_CHECK(read_flags(b, f, &active));
    while (active > 0) {
      B_HIP(hipGraphLaunch(b->g_chunk, b->stream));
      B_HIP(hipGraphLaunch(b->g_start1, b->stream));
      B_CHECK(read_flags(b, f, &active));
      ++chunks;
    }
    const double t2 = now_s();
    B_HIP(hipGraphLaunch(b->g_post, b->stream));
    B_HIP(hipMemcpyAsync(oc.data(), b->oc, sizeof(OcState) * b->np, hipMemcpyDeviceToHost,
                         b->stream));
    B_HIP(hipMemcpyAsync(compl_.data(), b->compl_, sizeof(double) * b->np,
                         hipMemcpyDeviceToHost, b->stream));
    B_HIP(hipStreamSynchronize(b->stream));
    const double t3 = now_s();
    if (hist) {
      for (int i = 0; i < b->np; ++i) {
        amt_iteration_result& R = hist[(size_t)it * b->np + i];
        std::memset(&R, 0, sizeof(R));
        R.compliance = compl_[i];
        const double nd = b->ndesign_h[i] > 0 ? b->ndesign_h[i] : (double)b->hp[i].g.n_elem;
        R.volume_fraction = oc[i].vol_sum / nd;
        R.max_density_change = oc[i].change;
        R.bisect_passes = oc[i].passes;
        R.solver_iterations = f[i].it;
        R.solver_converged = f[i].converged;
        R.outer_iterations = chunks;  // residual replacements
        R.true_rel_residual = f[i].rhs2 > 0 ? std::sqrt(f[i].rr / f[i].rhs2) : 0.0;
        R.assembly_time_s = t1 - t0;
        R.solve_time_s = t2 - t1;
        R.update_time_s = t3 - t2;
      }
    }
  }
  if (elapsed) *elapsed = now_s() - t_begin;
  return AMT_OK;
}

}  // namespace

// Multigrid mode: kBMgWorkers host threads take whole problems from a shared counter and run
// their `iterations` OC iterations on the problem's own handle and stream (single-problem
// semantics, multigrid FCG). Problems are independent, so the order in which they run does
// not change any result.
int run_mg(amt_batch* b, int iterations, amt_iteration_result* hist, double* elapsed_s) {
  for (int i = 0; i < b->np; ++i)
    if (!b->loads_set[i])
      return bfail(AMT_EINVAL, "problem " + std::to_string(i) + ": loads not set");
  const auto t0 = std::chrono::steady_clock::now();
  std::atomic<int> next{0}, failed{-1};
  std::vector<int> rcs(b->np, AMT_OK);
  auto work = [&]() {
    for (;;) {
      const int i = next.fetch_add(1);
      if (i >= b->np || failed.load() >= 0) return;
      amt_problem* h = b->mg[i];
      int rc = AMT_OK;
      if (!b->mg_begun[i]) {
        rc = amt_begin_run(h);
        b->mg_begun[i] = 1;
      }
      for (int it = 0; it < iterations && rc == AMT_OK; ++it) {
        amt_iteration_result r;
        rc = amt_run_iteration(h, &r);
        if (rc == AMT_OK && hist) hist[(size_t)it * b->np + i] = r;
      }
      if (rc == AMT_OK) rc = amt_sync(h);
      if (rc != AMT_OK) {
        rcs[i] = rc;
        int expect = -1;
        failed.compare_exchange_strong(expect, i);
      }
    }
  };
  const int nw = std::min(kBMgWorkers, b->np);
  std::vector<std::thread> pool;
  for (int w = 1; w < nw; ++w) pool.emplace_back(work);
  work();
  for (auto& t : pool) t.join();
  if (elapsed_s)
    *elapsed_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const int f = failed.load();
  if (f >= 0)
    return bfail(rcs[f], "problem " + std::to_string(f) + ": " + amt_last_error(b->mg[f]));
  return AMT_OK;
}

extern "C" {

int amt_batch_create(const amt_config* cfgs, int n, amt_batch** out) {
  if (!cfgs || !out || n <= 0) return bfail(AMT_EINVAL, "need at least one problem");
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    const amt_config& c = cfgs[i];
    if (c.nx <= 0 || c.ny <= 0 || c.nz <= 0)
      return bfail(AMT_EINVAL, "problem " + std::to_string(i) + ": grid dimensions must be positive");
    if (!(c.element_size > 0) || !(c.E > 0) || !(c.nu > -1.0 && c.nu < 0.5))
      return bfail(AMT_EINVAL, "problem " + std::to_string(i) + ": invalid element size or material");
    if (!(c.filter_radius >= 0) || c.filter_radius > 64)
      return bfail(AMT_EINVAL, "problem " + std::to_string(i) + ": filter radius must be in [0, 64]");
    const bool mgc = cfgs[0].precond == AMT_PRECOND_MULTIGRID;
    if (c.density_index != AMT_INDEX_CONSISTENT ||
        c.precond != (mgc ? AMT_PRECOND_MULTIGRID : AMT_PRECOND_JACOBI))
      return bfail(AMT_ENOTSUP, "batch: consistent density index and one solver for all "
                                "problems (Jacobi-PCG or multigrid)");
    if (c.precision != cfgs[0].precision || c.warm_start != cfgs[0].warm_start)
      return bfail(AMT_EINVAL, "batch: precision and warm_start must agree across problems");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return bfail(AMT_EHIP, "no HIP device available");
  if (cfgs[0].precond == AMT_PRECOND_MULTIGRID) {
    amt_batch* b = new amt_batch();
    b->np = n;
    b->cfg.assign(cfgs, cfgs + n);
    b->loads_set.assign(n, 0);
    b->mg_begun.assign(n, 0);
    for (int i = 0; i < n; ++i) {
      amt_problem* h = nullptr;
      if (amt_create(&cfgs[i], &h) != AMT_OK) {
        const std::string msg = amt_last_error(nullptr);
        destroy(b);
        return bfail(AMT_EHIP, "problem " + std::to_string(i) + ": " + msg);
      }
      b->mg.push_back(h);
    }
    *out = b;
    return AMT_OK;
  }
  for (int i = 0; i < n; ++i)
    if (cfgs[i].operator_tile < -1 || cfgs[i].operator_tile >= kBCands ||
        cfgs[i].operator_tile != cfgs[0].operator_tile)
      return bfail(AMT_EINVAL, "batch: operator_tile must be -1 (tuned) or a candidate 0.." +
                                   std::to_string(kBCands - 1) + ", the same for all problems");
  amt_batch* b = new amt_batch();
  b->np = n;
  b->mixed = cfgs[0].precision == AMT_PREC_MIXED_FP32;
  b->warm = cfgs[0].warm_start != 0;
  b->cfg.assign(cfgs, cfgs + n);
  b->loads_set.assign(n, 0);
  auto bail = [&](int rc) {
    destroy(b);
    return rc;
  };
  // descriptors, Ke0 tables (one per distinct E, nu, h) and stencils
  std::vector<double> ke64;
  std::vector<float> ke32;
  std::vector<double> ke_key;
  std::vector<StencilEntry> st_all;
  std::vector<double> sw_all;
  std::vector<int> ipv[BK_KINDS];
  for (int i = 0; i < n; ++i) {
    const amt_config& c = cfgs[i];
    BProb P;
    std::memset(&P, 0, sizeof(P));
    P.g.nx = c.nx;
    P.g.ny = c.ny;
    P.g.nz = c.nz;
    P.g.nny = c.ny + 1;
    P.g.nnz = c.nz + 1;
    P.g.n_nodes = (int64_t)(c.nx + 1) * (c.ny + 1) * (c.nz + 1);
    P.g.n_elem = (int64_t)c.nx * c.ny * c.nz;
    P.tg = make_tile_grid(P.g);
    P.e_off = b->n_e;
    P.n_off = b->n_n;
    b->n_e += P.g.n_elem;
    b->n_n += P.g.n_nodes;
    const int64_t cnt[BK_KINDS] = {P.tg.n_tiles, (P.g.n_elem + kBElemItem - 1) / kBElemItem,
                                   (3 * P.g.n_nodes + kBDofItem - 1) / kBDofItem,
                                   (P.g.n_nodes + kBlock - 1) / kBlock};
    for (int k = 0; k < BK_KINDS; ++k) {
      P.item0[k] = b->n_items[k];
      P.nitem[k] = (int)cnt[k];
      if ((int64_t)b->n_items[k] + cnt[k] > (int64_t)INT32_MAX / 64) {
        delete b;
        return bfail(AMT_EINVAL, "batch too large");
      }
      b->n_items[k] += (int)cnt[k];
      ipv[k].insert(ipv[k].end(), (size_t)cnt[k], i);
    }
    int ke = -1;
    for (size_t t = 0; t < ke_key.size() / 3; ++t)
      if (ke_key[3 * t] == c.E && ke_key[3 * t + 1] == c.nu && ke_key[3 * t + 2] == c.element_size)
        ke = (int)t;
    if (ke < 0) {
      ke = (int)(ke_key.size() / 3);
      ke_key.insert(ke_key.end(), {c.E, c.nu, c.element_size});
      double K[576];
      hex8_stiffness(c.E, c.nu, c.element_size, K);
      for (int q = 0; q < 576; ++q) {
        ke64.push_back(K[q]);
        ke32.push_back((float)K[q]);
      }
    }
    P.ke = ke;
    std::vector<StencilEntry> st;
    std::vector<double> sw;
    build_stencil(c.filter_radius, st, sw);
    P.st_off = (int)st_all.size();
    P.n_st = (int)st.size();
    st_all.insert(st_all.end(), st.begin(), st.end());
    sw_all.insert(sw_all.end(), sw.begin(), sw.end());
    P.maxit = c.pcg_maxit;
    P.penal = c.penalty;
    P.Emin = c.E_min;
    P.rho_min = c.rho_min;
    P.move = c.move_limit;
    P.vf = c.volume_fraction;
    P.pf = c.penalty_factor;
    P.tol = c.pcg_tol;
    b->hp.push_back(P);
  }
  b->n_dof = 3 * b->n_n;
  if (b->n_dof > (int64_t)INT32_MAX * 4) {
    delete b;
    return bfail(AMT_EINVAL, "batch too large");
  }
  // matvec candidates: the Hadamard blocks of every Ke0 table (the streaming candidates need
  // them), then per candidate each problem's items and the range of its partials
  std::vector<HadK<double>> hk64(ke_key.size() / 3);
  std::vector<HadK<float>> hk32(ke_key.size() / 3);
  b->had_ok = true;
  for (size_t t = 0; t < hk64.size(); ++t)
    b->had_ok = b->had_ok && hadamard_blocks<double>(&ke64[576 * t], hk64[t]) &&
                hadamard_blocks<float>(&ke64[576 * t], hk32[t]);
  if (cfgs[0].operator_tile > 0 && !b->had_ok) {
    delete b;
    return bfail(AMT_ENOTSUP, "batch: the streaming operator needs an isotropic Ke0");
  }
  std::vector<int> ipc[kBCands], arc[kBCands];
  std::vector<BStream> dsc[kBCands];
  int max_items = b->n_items[BK_TILE];
  for (int i = 0; i < n; ++i)
    arc[0].insert(arc[0].end(), {b->hp[i].item0[BK_TILE], b->hp[i].nitem[BK_TILE]});
  for (int c = 1; c < kBCands; ++c) {
    const int sh = (c - 1) / kBPlaneCounts, L = kBPlanes[(c - 1) % kBPlaneCounts];
    for (int i = 0; i < n; ++i) {
      BStream B;
      B.sg = make_stream_grid(b->hp[i].g, kBShapeTJ[sh], kBShapeTK[sh], L);
      B.item0 = b->nS[c];
      B.nitem = B.sg.n_tiles;
      if ((int64_t)b->nS[c] + B.nitem > (int64_t)INT32_MAX / 64) {
        delete b;
        return bfail(AMT_EINVAL, "batch too large");
      }
      b->nS[c] += B.nitem;
      ipc[c].insert(ipc[c].end(), (size_t)B.nitem, i);
      dsc[c].push_back(B);
      arc[c].insert(arc[c].end(), {B.item0, B.nitem});
    }
    max_items = std::max(max_items, b->nS[c]);
  }
  b->op_cand = cfgs[0].operator_tile;
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return bfail(AMT_EHIP, "stream creation failed");
  }
  const size_t tsz = b->mixed ? sizeof(float) : sizeof(double);
  const int64_t nd = b->n_dof, ne = b->n_e;
  int rc = AMT_OK;
  auto A = [&](int r) {
    if (rc == AMT_OK) rc = r;
  };
  A(balloc(b, &b->D, n));
  for (int k = 0; k < BK_KINDS; ++k) A(balloc(b, &b->ip[k], ipv[k].size()));
  A(balloc(b, &b->rho, ne));
  A(balloc(b, &b->s64, ne));
  A(balloc(b, &b->s32, ne));
  A(balloc(b, &b->dc, ne));
  A(balloc(b, &b->dcf, ne));
  A(balloc(b, &b->design, ne));
  A(balloc(b, &b->fixcnt, nd));
  A(balloc(b, &b->u, nd));
  A(balloc(b, &b->r, nd));
  A(balloc(b, &b->F, nd));
  A(balloc(b, &b->Fb, nd));
  A(balloc(b, &b->dinv, nd));
  A(balloc(b, &b->diag, nd));
  uint8_t *zz = nullptr, *pa = nullptr, *pb = nullptr, *qq = nullptr;
  A(balloc(b, &zz, nd * tsz));
  A(balloc(b, &pa, nd * tsz));
  A(balloc(b, &pb, nd * tsz));
  A(balloc(b, &qq, nd * tsz));
  b->z = zz;
  b->p0 = pa;
  b->p1 = pb;
  b->q = qq;
  A(balloc(b, &b->Ke64, ke64.size()));
  A(balloc(b, &b->Ke32, ke32.size()));
  A(balloc(b, &b->st, st_all.size()));
  A(balloc(b, &b->sw, sw_all.size()));
  A(balloc(b, &b->S, 2 * (size_t)n));
  A(balloc(b, &b->flags, n));
  A(balloc(b, &b->oc, n));
  A(balloc(b, &b->maxdiag, n));
  A(balloc(b, &b->bb, n));
  A(balloc(b, &b->pen, n));
  A(balloc(b, &b->ndesign, n));
  A(balloc(b, &b->compl_, n));
  A(balloc(b, &b->part_tile, 2 * (size_t)b->n_items[BK_TILE]));
  A(balloc(b, &b->part_dof, 2 * (size_t)b->n_items[BK_DOF]));
  A(balloc(b, &b->part_elem, (size_t)kBOcCand * b->n_items[BK_ELEM]));
  A(balloc(b, &b->part_node, b->n_items[BK_NODE]));
  A(balloc(b, &b->part_apply, (size_t)max_items));
  A(balloc(b, &b->HK64, hk64.size()));
  A(balloc(b, &b->HK32, hk32.size()));
  for (int c = 0; c < kBCands; ++c) {
    A(balloc(b, &b->ar[c], arc[c].size()));
    if (c > 0) {
      A(balloc(b, &b->ipS[c], ipc[c].size()));
      A(balloc(b, &b->DS[c], dsc[c].size()));
    }
  }
  if (rc != AMT_OK) return bail(rc);
  auto up = [&](void* dst, const void* src, size_t bytes) {
    if (rc == AMT_OK && bytes &&
        hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice) != hipSuccess)
      rc = bfail(AMT_EHIP, "upload failed");
  };
  up(b->D, b->hp.data(), sizeof(BProb) * n);
  for (int k = 0; k < BK_KINDS; ++k) up(b->ip[k], ipv[k].data(), sizeof(int) * ipv[k].size());
  up(b->Ke64, ke64.data(), sizeof(double) * ke64.size());
  up(b->Ke32, ke32.data(), sizeof(float) * ke32.size());
  up(b->st, st_all.data(), sizeof(StencilEntry) * st_all.size());
  up(b->sw, sw_all.data(), sizeof(double) * sw_all.size());
  up(b->HK64, hk64.data(), sizeof(HadK<double>) * hk64.size());
  up(b->HK32, hk32.data(), sizeof(HadK<float>) * hk32.size());
  for (int c = 0; c < kBCands; ++c) {
    up(b->ar[c], arc[c].data(), sizeof(int) * arc[c].size());
    if (c > 0) {
      up(b->ipS[c], ipc[c].data(), sizeof(int) * ipc[c].size());
      up(b->DS[c], dsc[c].data(), sizeof(BStream) * dsc[c].size());
    }
  }
  // initial density: each problem's volume fraction (topology_optimizer.cpp:403-405)
  {
    std::vector<double> rho(ne);
    for (int i = 0; i < n; ++i)
      std::fill(rho.begin() + b->hp[i].e_off, rho.begin() + b->hp[i].e_off + b->hp[i].g.n_elem,
                cfgs[i].volume_fraction);
    up(b->rho, rho.data(), sizeof(double) * ne);
  }
  if (rc != AMT_OK) return bail(rc);
  if (hipMemset(b->fixcnt, 0, nd) != hipSuccess || hipMemset(b->F, 0, nd * 8) != hipSuccess ||
      hipMemset(b->u, 0, nd * 8) != hipSuccess || hipDeviceSynchronize() != hipSuccess)
    return bail(bfail(AMT_EHIP, "initialisation failed"));
  *out = b;
  return AMT_OK;
}

int amt_batch_destroy(amt_batch* b) {
  destroy(b);
  return AMT_OK;
}

int amt_batch_size(const amt_batch* b, int i, int64_t* n_elements, int64_t* n_dofs) {
  if (!b || i < 0 || i >= b->np) return bfail(AMT_ERANGE, "problem index out of range");
  if (!b->mg.empty()) {
    const amt_config& c = b->cfg[i];
    if (n_elements) *n_elements = (int64_t)c.nx * c.ny * c.nz;
    if (n_dofs) *n_dofs = 3LL * (c.nx + 1) * (c.ny + 1) * (c.nz + 1);
    return AMT_OK;
  }
  if (n_elements) *n_elements = b->hp[i].g.n_elem;
  if (n_dofs) *n_dofs = 3 * b->hp[i].g.n_nodes;
  return AMT_OK;
}

int amt_batch_set_loads(amt_batch* b, int i, const double* F, int64_t n_dof,
                        const int64_t* fixed, int64_t n_fixed) {
  if (!b || i < 0 || i >= b->np) return bfail(AMT_ERANGE, "problem index out of range");
  if (!b->mg.empty()) {
    const int rc = amt_set_loads(b->mg[i], F, n_dof, fixed, n_fixed);
    if (rc != AMT_OK) return bfail(rc, amt_last_error(b->mg[i]));
    b->loads_set[i] = 1;
    return AMT_OK;
  }
  const BProb& P = b->hp[i];
  const int64_t nd = 3 * P.g.n_nodes;
  if (!F || n_dof != nd)
    return bfail(AMT_EINVAL, "force vector size mismatch: expected " + std::to_string(nd) +
                                 ", got " + std::to_string(n_dof));
  if (n_fixed < 0 || (n_fixed > 0 && !fixed)) return bfail(AMT_EINVAL, "bad fixed list");
  std::vector<uint8_t> cnt(nd, 0);
  for (int64_t t = 0; t < n_fixed; ++t) {
    const int64_t d = fixed[t];
    if (d < 0 || d >= nd)
      return bfail(AMT_ERANGE, "Fixed DOF index out of range: " + std::to_string(d));
    if (cnt[d] == 255) return bfail(AMT_EINVAL, "DOF listed more than 255 times");
    cnt[d]++;
  }
  for (int64_t t = 0; t < nd; ++t)
    if (!std::isfinite(F[t])) return bfail(AMT_ENAN, "non-finite force");
  B_HIP(hipStreamSynchronize(b->stream));
  B_HIP(hipMemcpy(b->F + 3 * P.n_off, F, sizeof(double) * nd, hipMemcpyHostToDevice));
  B_HIP(hipMemcpy(b->fixcnt + 3 * P.n_off, cnt.data(), nd, hipMemcpyHostToDevice));
  B_HIP(hipMemset(b->u + 3 * P.n_off, 0, sizeof(double) * nd));
  b->loads_set[i] = 1;
  return AMT_OK;
}

int amt_batch_set_density(amt_batch* b, int i, const double* rho, int64_t n_e) {
  if (!b || i < 0 || i >= b->np) return bfail(AMT_ERANGE, "problem index out of range");
  if (!b->mg.empty()) {
    const int rc = amt_set_density(b->mg[i], rho, n_e);
    if (rc != AMT_OK) return bfail(rc, amt_last_error(b->mg[i]));
    return AMT_OK;
  }
  const BProb& P = b->hp[i];
  if (!rho || n_e != P.g.n_elem)
    return bfail(AMT_EINVAL, "Density vector size mismatch: expected " +
                                 std::to_string(P.g.n_elem) + ", got " + std::to_string(n_e));
  for (int64_t t = 0; t < n_e; ++t)
    if (!std::isfinite(rho[t])) return bfail(AMT_ENAN, "non-finite density");
  B_HIP(hipStreamSynchronize(b->stream));
  B_HIP(hipMemcpy(b->rho + P.e_off, rho, sizeof(double) * n_e, hipMemcpyHostToDevice));
  return AMT_OK;
}

int amt_batch_get_density(amt_batch* b, int i, double* rho, int64_t n_e) {
  if (!b || i < 0 || i >= b->np) return bfail(AMT_ERANGE, "problem index out of range");
  if (!b->mg.empty()) {
    const int rc = amt_get_density(b->mg[i], rho, n_e);
    if (rc != AMT_OK) return bfail(rc, amt_last_error(b->mg[i]));
    return AMT_OK;
  }
  const BProb& P = b->hp[i];
  if (!rho || n_e != P.g.n_elem) return bfail(AMT_EINVAL, "density buffer size mismatch");
  B_HIP(hipStreamSynchronize(b->stream));
  B_HIP(hipMemcpy(rho, b->rho + P.e_off, sizeof(double) * n_e, hipMemcpyDeviceToHost));
  return AMT_OK;
}

int amt_batch_get_displacement(amt_batch* b, int i, double* u, int64_t n_dof) {
  if (!b || i < 0 || i >= b->np) return bfail(AMT_ERANGE, "problem index out of range");
  if (!b->mg.empty()) {
    const int rc = amt_get_displacement(b->mg[i], u, n_dof);
    if (rc != AMT_OK) return bfail(rc, amt_last_error(b->mg[i]));
    return AMT_OK;
  }
  const BProb& P = b->hp[i];
  if (!u || n_dof != 3 * P.g.n_nodes) return bfail(AMT_EINVAL, "displacement buffer size mismatch");
  B_HIP(hipStreamSynchronize(b->stream));
  B_HIP(hipMemcpy(u, b->u + 3 * P.n_off, sizeof(double) * n_dof, hipMemcpyDeviceToHost));
  return AMT_OK;
}

int amt_batch_operator_info(amt_batch* b, int* candidate, double* us, int n) {
  if (!b) return bfail(AMT_EINVAL, "bad argument");
  if (!b->mg.empty()) return bfail(AMT_ENOTSUP, "a multigrid batch has no lockstep matvec");
  if (candidate) *candidate = b->op_cand;
  if (us)
    for (int c = 0; c < n && c < kBCands; ++c) us[c] = b->op_us[c];
  return AMT_OK;
}

int amt_batch_run(amt_batch* b, int iterations, amt_iteration_result* hist, double* elapsed_s) {
  if (!b || iterations < 0) return bfail(AMT_EINVAL, "bad argument");
  if (!b->mg.empty()) return run_mg(b, iterations, hist, elapsed_s);
  return b->mixed ? run_impl<float>(b, iterations, hist, elapsed_s)
                  : run_impl<double>(b, iterations, hist, elapsed_s);
}

}  // extern "C"
