/*
 * amt.h — C-ABI of the MI355X SIMP topology-optimization engine (libamt.so).
 *
 * Plain pointers and sizes only; no exceptions cross this boundary. Every entry point
 * returns AMT_OK or an AMT_E* status; amt_last_error() gives the message.
 *
 * One handle = one problem on one GPU. The handle owns every device buffer and one HIP
 * stream; calls on one handle are not thread-safe. Host buffers passed in are copied.
 *
 * The reference interface each entry point replaces is cited as
 * cpp_engine/<file>:<line> of ettoremiglioranza1012/AM_ML_suite.
 *
 * Index conventions (same as the reference):
 *   node id (z fastest) = i*(ny+1)*(nz+1) + j*(nz+1) + k; dof = 3*node + c
 *     (element_kernel.cpp:186-216)
 *   density slot e      = iz*nx*ny + iy*nx + ix (x fastest) for sensitivities, filter, OC
 *     (topology_optimizer.hpp:218-232)
 *   the stiffness of element (i,j,k) reads density[i*ny*nz + j*nz + k] when
 *   density_index == AMT_INDEX_REF_COMPAT (global_assembler.hpp:162-164, as the reference),
 *   or density[k*nx*ny + j*nx + i] when AMT_INDEX_CONSISTENT.
 */
#ifndef AMT_H
#define AMT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AMT_OK 0
#define AMT_EINVAL 1   /* std::invalid_argument in the reference */
#define AMT_ERANGE 2   /* std::out_of_range in the reference */
#define AMT_EHIP 3     /* HIP runtime error (no device, launch failure, OOM) */
#define AMT_ENAN 4     /* non-finite input */
#define AMT_ENOTSUP 5  /* option not available in this build */

enum { AMT_INDEX_CONSISTENT = 0, AMT_INDEX_REF_COMPAT = 1 };
/* Precision of the FP32 work. MIXED_FP32: FP32 vector operator. MIXED_BF16: the operator on
 * the matrix cores (element-batched Ke0 x_e with x and Ke0 split into two BF16 terms each,
 * FP32 accumulation, ~1e-5 relative). Results are always held to the FP64 tolerance. */
enum { AMT_PREC_FP64 = 0, AMT_PREC_MIXED_FP32 = 1, AMT_PREC_MIXED_BF16 = 2 };
/* Preconditioners. JACOBI: the reference's solve with Jacobi in place of IC (FP64, or FP32
 * inner solves in FP64 refinement with a MIXED precision). MULTIGRID: flexible CG with FP64
 * vectors, operator and reductions around an FP32 Galerkin geometric-multigrid V-cycle (its
 * fine-level smoothing on the matrix cores with AMT_PREC_MIXED_BF16). */
enum { AMT_PRECOND_JACOBI = 0, AMT_PRECOND_UNET = 1, AMT_PRECOND_MULTIGRID = 2 };

typedef struct amt_problem amt_problem;

/* Problem + SIMPConfig (topology_optimizer.hpp:49-59) + MaterialProperties
 * (element_kernel.hpp:61-87) + the solver settings the optimizer hard-codes
 * (topology_optimizer.cpp:35-40), in one flat struct. */
typedef struct amt_config {
  int nx, ny, nz;
  double element_size;
  double E, nu;
  double volume_fraction, penalty, filter_radius, move_limit, rho_min, E_min;
  int max_iterations;
  double convergence_tol;
  double pcg_tol;        /* relative residual |r|/|F_bc| */
  int pcg_maxit;
  double penalty_factor; /* P = penalty_factor * max|K_ii| */
  int density_index;     /* AMT_INDEX_* */
  int precision;         /* AMT_PREC_* */
  int precond;           /* AMT_PRECOND_* */
  int warm_start;        /* 1: start each solve from the previous displacement; 2 (multigrid
                            solver): from the energy-minimising combination of the previous two
                            displacements (one extra operator application per solve) */
  double inner_tol;      /* mixed precision: relative reduction per FP32 inner solve */
  int inner_maxit;
  int verbose;
  int use_graphs;        /* 1: replay the multigrid CG iteration chunk as a HIP graph */
  int operator_tile;     /* level-0 operator launch of the multigrid solver: -1 autotune on the
                            first solve (default); 0 the 4x4x16 node gather tile; 1 + 3 s + l the
                            element-column streaming tile of shape s (0: 16x16, 1: 8x32, 2: 4x64
                            threads along j x k) with 4, 8 or 16 (l = 0, 1, 2) node planes per
                            block; 10 + 2 s + l the same shapes with 2 or 3 (l = 0, 1) planes.
                            All streaming shapes give the same operator values. In a lockstep
                            batch (amt_batch_create): the PCG matvec, -1 tuned at the first
                            run, 0 the gather tile, 1 + 3 s + l shape s with 2, 4 or 8 planes
                            (amt_batch_operator_info); the same value for every problem. */
  int mg_fused;          /* multigrid V-cycle on the levels of at most 16,384 nodes: 1 (default)
                            two brick-tiled launches per level (restriction + pre-smoothing,
                            prolongation + post-smoothing) and one for the dense level; 0 one
                            launch per smoothing or transfer step. Same linear map, summed in
                            another order. */
  int mg_coarse_cheb;    /* smoother of the multigrid levels from 2 on: 0 damped Jacobi (as on
                            levels 0 and 1), k = 2..4 the degree-k Chebyshev polynomial in
                            M^-1 A, M = diag(sum_j |A_ij|) (default 3); on grids of more than
                            2^18 elements the two stencil levels above the dense one take
                            degree k + 1 */
} amt_config;

/* IterationResult (topology_optimizer.hpp:65-74) plus solver detail. */
typedef struct amt_iteration_result {
  double compliance, volume_fraction, max_density_change;
  int solver_iterations; /* PCG passes (Eigen count: matvecs - 1 when converged) */
  int solver_converged;
  int bisect_passes;
  int outer_iterations;  /* FP64 refinement steps (1 for plain FP64) */
  int solver_fallback;   /* 1: the learned preconditioner stalled, Jacobi-PCG finished */
  double true_rel_residual; /* |F_bc - K u| / |F_bc|, recomputed in FP64 */
  double assembly_time_s, solve_time_s, sensitivity_time_s, update_time_s;
} amt_iteration_result;

/* Fills the reference defaults: SIMPConfig{} (topology_optimizer.hpp:49-59), Ti6Al4V
 * (element_kernel.hpp:67-69), solver tol 1e-6 / maxit 5000 / penalty 1e8
 * (topology_optimizer.cpp:35-40), reference density index, FP64, Jacobi. */
void amt_config_default(amt_config* cfg);

/* ElementKernel::computeStiffness (element_kernel.cpp:155-179): hex8 Ke0 of a cube of edge
 * h, 24x24 row-major. Host computation; needs no device. */
int amt_element_stiffness(double E, double nu, double h, double* Ke);

/* TopologyOptimizer ctor (topology_optimizer.cpp:16-53): Ke0, filter stencil, buffers. */
int amt_create(const amt_config* cfg, amt_problem** out);
int amt_destroy(amt_problem* p);
/* Message of the last failed call on p (p may be NULL for amt_create failures). */
const char* amt_last_error(const amt_problem* p);

/* forces (n_dof) and fixed DOF list, as passed to run()/runIteration()
 * (topology_optimizer.hpp:149-169). Out-of-range DOF -> AMT_ERANGE
 * (linear_solver.cpp:69-73). Duplicated DOFs add the penalty once per occurrence. */
int amt_set_loads(amt_problem* p, const double* F, int64_t n_dof, const int64_t* fixed,
                  int64_t n_fixed);
int amt_set_density(amt_problem* p, const double* rho, int64_t n_e);
int amt_get_density(amt_problem* p, double* rho, int64_t n_e);
int amt_get_displacement(amt_problem* p, double* u, int64_t n_dof);

/* run() prologue (topology_optimizer.cpp:419-428): design mask from the current density. */
int amt_begin_run(amt_problem* p);

/* runIteration (topology_optimizer.cpp:294-384) on the device-resident density. If no
 * design mask exists yet it is built from the current density first. */
int amt_run_iteration(amt_problem* p, amt_iteration_result* out);

/* IterationCallback (topology_optimizer.hpp:98-102); return 0 to stop. */
typedef int (*amt_iter_cb)(int iteration, const double* density, int64_t n_e,
                           const amt_iteration_result* result, void* user);

/* run (topology_optimizer.cpp:391-520) from the density set by amt_set_density.
 * hist must hold max_it entries. */
int amt_run(amt_problem* p, int max_it, amt_iteration_result* hist, int* iterations,
            int* converged, double* elapsed_s, amt_iter_cb cb, void* user);

/* ---- phase entry points (the reference's private phases; used by parity tests) ---- */

/* y = K x with the current density (global_assembler.cpp:61-222 + SpMV). with_bc=1 adds
 * the penalty rows (linear_solver.cpp:35-91) exactly as the solve sees them. */
int amt_apply_operator(amt_problem* p, const double* x, double* y, int with_bc);
/* As amt_apply_operator with the operator of a given precision (AMT_PREC_*): FP64 vector,
 * FP32 vector, or split-BF16 MFMA; x and y are FP64 on the host either way. */
int amt_apply_operator_prec(amt_problem* p, const double* x, double* y, int with_bc,
                            int precision);
/* Diagonal of K (+ penalty if with_bc) and P = penalty_factor * max|K_ii|. */
int amt_get_diagonal(amt_problem* p, double* diag, int with_bc, double* penalty);

typedef struct amt_solve_stats {
  int iterations, converged, outer_iterations, fallback;
  double rel_residual, true_rel_residual, solve_time_s;
} amt_solve_stats;
/* solveWithBCs (linear_solver.cpp:186-203) for the current density and loads; u (n_dof)
 * receives the displacement. */
int amt_solve(amt_problem* p, double* u, amt_solve_stats* stats);

/* One application z = M r of the multigrid preconditioner (handles with precond
 * AMT_PRECOND_MULTIGRID) for the current density and loads; r and z have n_dof entries. No
 * counterpart in the reference (its IncompleteCholesky preconditioner is internal to Eigen's
 * ConjugateGradient, linear_solver.cpp:133-140); it exposes the V-cycle for tests. */
int amt_apply_multigrid(amt_problem* p, const double* r, double* z);

/* computeSensitivities (topology_optimizer.cpp:126-161) for u (n_dof) and the current
 * density; dc has n_e entries (slot order). */
int amt_sensitivities(amt_problem* p, const double* u, double* dc);
/* filterSensitivities (topology_optimizer.cpp:168-198) with the current density. */
int amt_filter(amt_problem* p, const double* dc, double* dcf);
/* Neighbour list of slot e as the filter kernel walks it (topology_optimizer.cpp:60-119).
 * Writes up to cap entries; *count = list length. */
int amt_filter_neighbors(amt_problem* p, int64_t e, int64_t* idx, double* w, int cap,
                         int* count);
/* updateDensitiesOC (topology_optimizer.cpp:205-287) on the device density with dcf (n_e)
 * and the current design mask. */
int amt_oc_update(amt_problem* p, const double* dcf, double* max_change, int* passes);

/* ---- learned preconditioner (BF16 3D U-Net; no counterpart in the reference, whose
 *      src/am/ai/model.py:48-51 raises NotImplementedError) ----
 * M(r) = D^-1 r + gain * rms * net(density feature, D^-1 r / rms) on free rows (D^-1 r on
 * penalty rows) inside flexible CG, with plain Jacobi-PCG taking over when |r| stagnates
 * (16 four-iteration chunks without a 10% drop) or the iteration breaks down. With warm_start
 * and no previous displacement, u0 = gamma M(F_bc), gamma energy-optimal. A new network has
 * gain 0 (plain Jacobi); training sets it. Networks are fully convolutional: parameters and
 * gain move between grids with get/set_params and get/set_gain. */
typedef struct amt_unet amt_unet;
/* Network for a padded voxel grid (each dimension a positive multiple of 8). */
int amt_unet_create(int px, int py, int pz, uint64_t seed, amt_unet** out);
/* Network sized for p's node grid padded to multiples of 8. */
int amt_unet_create_for(amt_problem* p, uint64_t seed, amt_unet** out);
int amt_unet_destroy(amt_unet* net);
int64_t amt_unet_num_params(const amt_unet* net);
int amt_unet_get_params(amt_unet* net, float* params, int64_t n);
int amt_unet_set_params(amt_unet* net, const float* params, int64_t n);
/* Forward on host tensors: x, y are [px*py*pz][16] FP32 (stored BF16 on the device). */
int amt_unet_forward(amt_unet* net, const float* x, float* y);
/* The forward pass of inference (the preconditioner's): the same output as amt_unet_forward,
 * without writing the training intermediates (the decoder reads its upsampled inputs in
 * place), so amt_unet_backward / amt_unet_activation of the decoder's upsampled tensors refer to
 * the last amt_unet_forward. */
int amt_unet_infer(amt_unet* net, const float* x, float* y);
/* Times reps inference forward passes on the network's stored input (after one untimed pass)
 * with HIP events on its stream: *ms_per_forward, and the convolutions' multiply-add rate over the whole
 * pass (*conv_tflops, 2 x taps x cin x cout per output voxel of each layer, pooling and
 * upsampling time included). */
int amt_unet_benchmark(amt_unet* net, int reps, double* ms_per_forward, double* conv_tflops);
/* Parameter gradient of sum(y * dy) for the last forward; dparams has num_params entries. */
int amt_unet_backward(amt_unet* net, const float* dy, float* dparams);
/* Copy a stored activation of the last forward to the host as FP32. tensor indexes
 * X0 A0a S0 P1 A1a S1 P2 A2a S2 P3 B3a B3 U2 C2a C2 U1 C1a C1 U0 C0a C0 OUT (0..21), the
 * network's tensors in forward order; n = voxels of its level * its channel count. */
int amt_unet_activation(amt_unet* net, int tensor, float* out, int64_t n);
/* Attach (or detach with NULL) a network. With AMT_PRECOND_UNET it is the preconditioner above.
 * With AMT_PRECOND_MULTIGRID and warm_start it is a learned warm start: every solve that starts
 * from the previous displacement u first takes the step u += gamma M(r), r = F_bc - K u,
 * gamma = r.M(r) / M(r).K M(r) (energy-optimal; 0 if the curvature is not positive), then runs
 * the flexible CG from there, held to the same FP64 tolerance. */
int amt_attach_unet(amt_problem* p, amt_unet* net);
/* Appends training samples: residuals of a plain Jacobi-PCG solve of p's current problem
 * (every 32 passes, up to max_samples). *n_total = samples held. */
int amt_unet_collect(amt_problem* p, int max_samples, int* n_total);
/* Arms warm-start sample collection on a multigrid handle: while fewer than max_samples are
 * held, every solve that starts from the previous displacement appends its start residual
 * F_bc - K u (with the density it belongs to) to the samples amt_unet_train cycles through. */
int amt_unet_collect_warm(amt_problem* p, int max_samples, int* n_total);
/* Adam steps on the held samples (cycled in order; each sample is taken with the operator and
 * density feature of its own density, so the handle's density is the last sample's afterwards). Each step forms U = D^-1 r + theta N
 * (N the network's correction, theta the energy-optimal weight, folded into the network's gain
 * as a running mean) and descends the energy (U.KU - 2 U.r) / (r.D^-1 r); two FP32 operator
 * applications per step. losses / jacobi_losses (optional, steps entries) get each step's
 * energy and that of U = D^-1 r alone. lr = 0 evaluates: parameters and gain stay as they are. */
int amt_unet_train(amt_problem* p, amt_unet* net, int steps, double lr, double* losses,
                   double* jacobi_losses);
/* Weight of the correction used by the preconditioner (0 for a new network). */
int amt_unet_get_gain(amt_unet* net, double* gain);
int amt_unet_set_gain(amt_unet* net, double gain);

/* ---- batched dataset generation (no counterpart in the reference, which optimises one
 *      problem per TopologyOptimizer, topology_optimizer.cpp:391-520) ----
 * n independent problems of different sizes, each with its own amt_config (grid, material,
 * volume fraction, filter radius, ...), advanced in lockstep: every phase is one persistent
 * launch over all problems, and the OC step is replayed from hipGraphs. Each problem follows
 * the single-problem semantics with the consistent density index and Jacobi-PCG:
 * precision AMT_PREC_MIXED_FP32 runs the PCG matvec in FP32 with FP64 displacement, residual
 * and reductions, replacing the recursive residual by the FP64 one every 16 iterations, until
 * |F_bc - K u| <= pcg_tol |F_bc|;
 * AMT_PREC_FP64 is plain FP64 Jacobi-PCG. precision and warm_start must agree across
 * problems. With precond = AMT_PRECOND_MULTIGRID for every problem, each problem instead gets
 * its own single-problem handle and HIP stream, and host threads run the problems' OC
 * iterations concurrently (results bitwise those of each problem run alone). Errors of these
 * calls are read with amt_last_error(NULL). */
typedef struct amt_batch amt_batch;
int amt_batch_create(const amt_config* configs, int n, amt_batch** out);
int amt_batch_destroy(amt_batch* b);
/* Element and DOF counts of problem i. */
int amt_batch_size(const amt_batch* b, int i, int64_t* n_elements, int64_t* n_dofs);
/* As amt_set_loads / amt_set_density / amt_get_density / amt_get_displacement, for problem i.
 * Densities default to each problem's volume fraction. */
int amt_batch_set_loads(amt_batch* b, int i, const double* F, int64_t n_dof,
                        const int64_t* fixed, int64_t n_fixed);
int amt_batch_set_density(amt_batch* b, int i, const double* rho, int64_t n_e);
int amt_batch_get_density(amt_batch* b, int i, double* rho, int64_t n_e);
int amt_batch_get_displacement(amt_batch* b, int i, double* u, int64_t n_dof);
/* `iterations` OC iterations of every problem (design masks from the densities at the first
 * call). hist (optional) gets iterations * n results, hist[it * n + i] for problem i;
 * solver_iterations counts PCG iterations, outer_iterations the residual replacements,
 * and the phase times are the batch's. elapsed_s (optional): wall time of the call. */
int amt_batch_run(amt_batch* b, int iterations, amt_iteration_result* hist, double* elapsed_s);
/* The lockstep PCG matvec: *candidate (0 the LDS gather tile; 1 + 3 s + l the element-column
 * streaming tile of shape s (0: 16x16, 1: 8x32, 2: 4x64 threads along j x k) with 2, 4 or 8
 * (l = 0, 1, 2) node planes), chosen at the first amt_batch_run as the fastest of the launches
 * timed there unless amt_config.operator_tile forced it (-1 before); us[c] (up to n entries,
 * optional) the measured launch time of candidate c in microseconds, 0 where it did not run.
 * A multigrid batch has no lockstep matvec (AMT_ENOTSUP). No counterpart in the reference. */
int amt_batch_operator_info(amt_batch* b, int* candidate, double* us, int n);

/* ---- The Python SIMPOptimizer flavour (src/am/numerical/topopt.py:84-411) ----
 * The suite's second SIMP loop, on the same matrix-free operator. It differs from
 * TopologyOptimizer in every phase around the solve:
 *   - design variable rho (nx, ny, nz) C-order; each iteration filters it,
 *     rho~ = clip(convolve(rho, w, mode='reflect'), rho_min, 1) (:154-156, :302-303), with
 *     w = cone kernel r - dist on the (2 ceil(r) + 1)^3 box, normalised to sum 1 (:134-151);
 *   - element scale s_e = E_eff / E: void 0 (skipped), fixed 1, design
 *     (E_min + rho~^p (E - E_min)) / E (fem.py:181-207);
 *   - per load case, K_ff u_f = F_f with the case's constrained DOFs and the orphan DOFs
 *     (|K_ii| < 1e-15) eliminated (fem.py:262-304; the reference solves directly, here FP64
 *     Jacobi-PCG to solve_tol), compliance sum_e E_ratio_e u_e^T Ke0 u_e (fem.py:356-384),
 *     dc_e += -p rho~_e^(p-1) u_e^T Ke0 u_e on design elements (topopt.py:158-198);
 *   - dcf = convolve(dc * rho, w, 'reflect') / max(rho, 1e-3) (:363);
 *   - OC bisection on [1e-9, 1e9], stop when (l2-l1)/max(l1+l2, 1e-10) <= 1e-3, Be =
 *     sqrt(-dcf / (lmid + 1e-10)), density of the last evaluated lmid (:200-262);
 *   - change = max |rho_new - rho|; volume = mean rho_new over the design elements. */
typedef struct amt_simp amt_simp;
typedef struct {
  double volume_fraction; /* SIMPParams defaults: 0.25 */
  double penalty;         /* 3.0 */
  double filter_radius;   /* 2.0 (elements) */
  double move_limit;      /* 0.2 */
  double rho_min;         /* 0.01 */
  double convergence_tol; /* 0.01 */
  double E_min;           /* 1e-9 */
  int max_iterations;     /* 100 */
  double solve_tol;       /* relative residual of the per-load-case solve, e.g. 1e-10 */
  int solve_maxit;
} amt_simp_params;
typedef struct {
  double compliance; /* sum over load cases */
  double volume;     /* mean density over design elements */
  double change;     /* max |rho_new - rho| */
  int solver_iterations;      /* summed over load cases */
  double max_true_rel_residual; /* worst load case, |M(F - K u)| / |M F| */
  int solver_converged;       /* all load cases reached solve_tol */
} amt_simp_iteration;
/* grid: (nx, ny, nz) C-order int8 (-1 void, 0 design, 1 fixed). Density starts at
 * volume_fraction everywhere (topopt.py:124). */
int amt_simp_create(int nx, int ny, int nz, double element_size, double E, double nu,
                    const int8_t* grid, const amt_simp_params* prm, amt_simp** out);
int amt_simp_destroy(amt_simp* h);
/* Appends a load case: F[3 (nx+1)(ny+1)(nz+1)] and its constrained DOFs (any order). */
int amt_simp_add_load_case(amt_simp* h, const double* F, int64_t n_dof,
                           const int64_t* constrained, int64_t n_constrained);
int amt_simp_set_density(amt_simp* h, const double* rho, int64_t n_e);
int amt_simp_get_density(amt_simp* h, double* rho, int64_t n_e);
/* One iteration of SIMPOptimizer.run's loop (topopt.py:294-383). */
int amt_simp_step(amt_simp* h, amt_simp_iteration* out);
/* y = convolve(x, w, mode='reflect') on the element grid. */
int amt_simp_filter(amt_simp* h, const double* x, double* y, int64_t n_e);
/* Displacement of load case lc from the last step. */
int amt_simp_get_displacement(amt_simp* h, int lc, double* u, int64_t n_dof);

/* ---- OC update without allocation ----
 * No counterpart in the reference. For callers that cannot allocate device memory: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc. The same update and
 * kernels as amt_oc_update (updateDensitiesOC, topology_optimizer.cpp:205-287), on caller
 * device buffers and a caller stream (NULL: the null stream), enqueued without a host sync:
 *   rho[n_e]    device, in/out: densities (slot order), updated in place;
 *   dcf[n_e]    device: filtered sensitivities;
 *   design[n_e] device: 1 for design elements (the mask amt_begin_run builds: rho_min + 1e-6 <
 *               rho < 0.99 at the run's start); an all-zero mask counts every element;
 *   result[4]   device, out: max |drho|, design volume sum, bisection passes, design count.
 * Scratch (partials and bisection state, ~266 KB) is one static device array per GPU, held
 * until the process ends; nothing is allocated. n_e is limited to AMT_OC_NOALLOC_MAX_ELEMENTS
 * (larger: AMT_EINVAL). Every call on one GPU shares that array, so calls that overlap on one
 * GPU (other streams or threads) race; calls on one stream are ordered. Errors: amt_last_error
 * (NULL). */
#define AMT_OC_NOALLOC_MAX_ELEMENTS ((int64_t)1 << 31)
int amt_oc_update_noalloc(double* rho, const double* dcf, const uint8_t* design, int64_t n_e,
                          double volume_fraction, double move_limit, double rho_min,
                          double* result, void* stream);

/* Wait for all work of this handle. */
int amt_sync(amt_problem* p);
/* Device name / CU count of the handle's GPU. */
int amt_device_info(char* name, int cap, int* n_cu);

/* The level-0 operator launch chosen for kind 0 (FP64 CG operator and residual), 1 (FP32
 * pre-smoothing with the flexible-CG update) or 2 (FP32 post-smoothing): *candidate as in
 * amt_config.operator_tile, the tile shape (*tj x *tk threads, *planes node planes; 4 x 16 x 4
 * node tile for candidate 0) and the launch time the autotuner measured (*us, 0 if it did not
 * run). Valid after the first multigrid solve (the tuner runs there); no counterpart in the
 * reference. */
int amt_operator_info(amt_problem* p, int kind, int* candidate, int* tj, int* tk, int* planes,
                      double* us);
/* The autotuner's measured launch time of every candidate for that kind, us[c] for candidate c
 * (as amt_config.operator_tile), up to n entries; 0 where it did not run. */
int amt_operator_timings(amt_problem* p, int kind, double* us, int n);

/* Kernel timing of the operator (K x) launches on the handle's stream: sampled launches go
 * through hipExtLaunchKernelGGL, whose events the GPU timestamps at the kernel's start and end.
 * Enable, then read the number of timed launches and their summed time. */
int amt_profile_enable(amt_problem* p, int on);
int amt_profile_read(amt_problem* p, int64_t* launches, double* total_ms);
/* As amt_profile_read for kind 0 (the operator launches: every 8th of the Jacobi-PCG solver's,
 * the flexible-CG operator's of the first chunk of every multigrid solve) or kind 1 (the
 * level-0 pre-smoothing with the flexible-CG update, first chunk of every multigrid solve).
 * No counterpart in the reference. */
int amt_profile_read_kind(amt_problem* p, int kind, int64_t* launches, double* total_ms);

#ifdef __cplusplus
}
#endif
#endif
