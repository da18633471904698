/*
 * oracle.h — CPU FP64 restatement of the reference SIMP topology-optimization
 * path (ettoremiglioranza1012/AM_ML_suite, cpp_engine/).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product
 * library (libamt.so) and the timed CPU baseline of bench.py. Nothing in the
 * product links, loads or calls it. It is compiled with -ffp-contract=off so
 * that every a*b+c rounds twice, as written.
 *
 * Index conventions (reference):
 *   node id  (z fastest)  = i*(ny+1)*(nz+1) + j*(nz+1) + k     element_kernel.cpp:207
 *   dof                   = 3*node + c                          element_kernel.cpp:210-212
 *   assembler density idx = i*ny*nz + j*nz + k  (C order)       global_assembler.hpp:162-164
 *   optimizer slot idx    = iz*nx*ny + iy*nx + ix (x fastest)   topology_optimizer.hpp:218-232
 */
#ifndef AM_ORACLE_H
#define AM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Ke0 by 2x2x2 Gauss quadrature of B^T C B |J| (element_kernel.cpp:71-179). Row-major 24x24. */
void orc_ke0(double E, double nu, double h, double* Ke /*576*/);

/* 24 global DOFs of element (i,j,k) (element_kernel.cpp:186-216). */
void orc_element_dofs(int i, int j, int k, int nx, int ny, int nz, int64_t* dofs /*24*/);

/* Assembled CSR K (global_assembler.cpp:61-222): triplets per element, |v|>1e-20 kept,
 * duplicates summed. density_index: 1 = reference (C-order read), 0 = consistent (x-fastest).
 * Call with rowptr!=NULL, col/val==NULL to get nnz only; returns nnz or -1 on error. */
int64_t orc_assemble_csr(int nx, int ny, int nz, double h, double E, double nu,
                         const double* rho, double penal, double Emin, double rho_min,
                         int ref_compat, int64_t* rowptr, int32_t* col, double* val);

/* y = K x for a CSR matrix. */
void orc_csr_spmv(int64_t n, const int64_t* rowptr, const int32_t* col, const double* val,
                  const double* x, double* y);

/* Penalty BCs (linear_solver.cpp:17-91): P = pf*max|K_ii|, K_ii += P, F_i = 0 for each
 * listed DOF (duplicates add twice, as coeffRef does). Returns P, or -1 on out-of-range DOF. */
double orc_apply_penalty(int64_t n, const int64_t* rowptr, const int32_t* col, double* val,
                         double* F, const int64_t* fixed, int64_t nfixed, double penalty_factor);

/* Jacobi-preconditioned CG from x=0 with Eigen's stopping rule
 * (|r|^2 < max(tol^2 |b|^2, DBL_MIN), iterations = completed loop passes).
 * Returns 1 if converged. true_rel = |b - K x| / |b| recomputed in FP64. */
int orc_pcg_jacobi(int64_t n, const int64_t* rowptr, const int32_t* col, const double* val,
                   const double* b, double* x, double tol, int maxit,
                   int* iters, double* rel_err, double* true_rel);

/* Matrix-free y = sum_e s_e Ke0 x_e over the grid (no BCs); s in C-order (i*ny*nz+j*nz+k). */
void orc_matfree_apply(int nx, int ny, int nz, const double* Ke, const double* s_c,
                       const double* x, double* y);

/* dc[e] = -p rho[e]^(p-1) u_e^T Ke0 u_e for slot e decoded x-fastest
 * (topology_optimizer.cpp:126-161). */
void orc_sensitivities(int nx, int ny, int nz, const double* Ke, const double* u,
                       const double* rho, double penal, double* dc);

/* Filter neighbour lists (topology_optimizer.cpp:60-119). offs has n_e+1 entries; idx/w may be
 * NULL to count only. Returns total entries. */
int64_t orc_build_filter(int nx, int ny, int nz, double rmin, int64_t* offs, int64_t* idx,
                         double* w);

/* dc_f[e] = sum w rho dc / sum w rho, fallback dc[e] if denominator <= 1e-15
 * (topology_optimizer.cpp:168-198). */
void orc_filter(int64_t n_e, const int64_t* offs, const int64_t* idx, const double* w,
                const double* dc, const double* rho, double* dcf);

/* OC bisection (topology_optimizer.cpp:205-287). design: 0/1 per slot. rho updated in place.
 * Returns max|drho|; n_pass/lmid_out report the bisection passes run and the last lambda. */
double orc_oc_update(int64_t n_e, double* rho, const double* dcf, const uint8_t* design,
                     int64_t n_design, double vf, double move, double rho_min,
                     int* n_pass, double* lmid_out);

/* Full optimizer run (topology_optimizer.cpp:16-53, 294-520). */
typedef struct {
  int nx, ny, nz;
  double h, E, nu;
  double volume_fraction, penalty, filter_radius, move_limit, rho_min, E_min;
  int max_iterations;
  double convergence_tol;
  double pcg_tol;
  int pcg_maxit;
  double penalty_factor;
  int ref_compat;  /* 1: assembler reads density C-order (reference); 0: x-fastest */
  int verbose;
} orc_config;

typedef struct {
  double compliance, volume_fraction, max_density_change;
  int solver_iterations, solver_converged, bisect_passes;
  double true_rel_residual;
  double assembly_time_s, solve_time_s, sensitivity_time_s, update_time_s;
} orc_iter_result;

/* Runs up to max_it OC iterations; rho (n_e) in/out; per-iteration results in res[max_it].
 * Returns iterations performed; *converged set as the reference does. u_out (optional)
 * receives the last displacement. */
int orc_run(const orc_config* cfg, double* rho, const double* F, const int64_t* fixed,
            int64_t nfixed, int max_it, orc_iter_result* res, int* converged, double* u_out);

/* BRK-A-01 voxel domain of am_topopt.cpp:98-239 (x-fastest grid, -1/0/1). Returns n_e and
 * writes nx,ny,nz; grid may be NULL for a size query. */
int64_t orc_bracket_domain(double res_mm, int8_t* grid, int* nx, int* ny, int* nz);

/* Eyelet ring load (am_topopt.cpp:248-314) into F (n_dof, zeroed here); returns node count. */
int orc_bracket_forces(int nx, int ny, int nz, double res_mm, double load_N, double* F);

/* Base clamp (am_topopt.cpp:319-334); returns count written to fixed (may be NULL). */
int64_t orc_bracket_fixed(int nx, int ny, int nz, int64_t* fixed);

int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
