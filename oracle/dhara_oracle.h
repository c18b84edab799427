/*
 * dhara_oracle.h -- CPU FP64 restatement of Dhara's GPU subsurface + overland path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load this library, and only as the checker (or as the timed CPU
 * baseline).  The product path (libdhara_hip.so) never links or calls it.
 *
 * It mirrors the dhara_* C-ABI of include/dhara_hip.h with a dor_ prefix, so the same
 * driver code runs against both.  On top of that it exposes the pieces the known-answer
 * tests need: CUSP-semantics DIA generation / SpMV / CG / BiCGStab / monitor, a dense LU
 * solve, and an independent 1D column Richards solver (Thomas algorithm).
 *
 * Pinning status of every function is listed in the header comment of dhara_oracle.cpp.
 */
#ifndef DHARA_ORACLE_H
#define DHARA_ORACLE_H

#include "../include/dhara_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Subsurface linear solver of the oracle. */
#define DOR_LIN_PCG_JACOBI   0  /* Jacobi-PCG on the assembled DIA matrix (tight mode)      */
#define DOR_LIN_BICGSTAB_REF 1  /* CUSP bicgstab, identity preconditioner
                                   (cusp/krylov/detail/bicgstab.inl:64-146)                */
#define DOR_LIN_DENSE        2  /* dense LU (tiny grids only)                               */
#define DOR_LIN_PCG_LINE     3  /* PCG with the column (vertical line) preconditioner        */
/* Overland linear solver of the oracle. */
#define DOR_OLF_PCG_JACOBI   0  /* Dirichlet ring eliminated, Jacobi-PCG (tight mode)       */
#define DOR_OLF_CG_REF       1  /* CUSP cg on the full ring-identity DIA matrix
                                   (cusp/krylov/detail/cg.inl:64-128)                       */
#define DOR_OLF_DENSE        2

typedef struct dor_ctx dor_ctx;

int dor_create(const dhara_grid* grid, const dhara_params* params, int32_t lin_solver,
               int32_t olf_solver, dor_ctx** out);
int dor_destroy(dor_ctx* ctx);
const char* dor_last_error(const dor_ctx* ctx);
int dor_set_params(dor_ctx* ctx, const dhara_params* params);
int dor_upload_field(dor_ctx* ctx, int32_t field, const double* host, size_t n);
int dor_upload_field_i32(dor_ctx* ctx, int32_t field, const int32_t* host, size_t n);
int dor_download_field(dor_ctx* ctx, int32_t field, double* host, size_t n);
int dor_download_field_i32(dor_ctx* ctx, int32_t field, int32_t* host, size_t n);
int dor_scatter_group_field(dor_ctx* ctx, int32_t field, const double* values, size_t ngroups);
int dor_init_state(dor_ctx* ctx, int32_t mode);
int dor_subsurface_step(dor_ctx* ctx, dhara_step_stats* stats);
int dor_overland_step(dor_ctx* ctx, dhara_step_stats* stats);
int dor_step(dor_ctx* ctx, int32_t nsteps, dhara_step_stats* stats);
int dor_eval_constitutive(dor_ctx* ctx, const double* psi, double* theta, double* K,
                          double* C);
int dor_eval_inverse(dor_ctx* ctx, const double* theta, double* psi);
/* Reference-form (pow-based) constitutive evaluation, src/global.cu:66-105 in SI. */
int dor_eval_constitutive_refform(dor_ctx* ctx, const double* psi, double* theta,
                                  double* K, double* C);
/* a3d: 7*N DIA values (may be NULL), diag/rhs: N (may be NULL). */
int dor_assemble_subsurface(dor_ctx* ctx, double* a3d, double* diag, double* rhs);
int dor_apply_subsurface_operator(dor_ctx* ctx, const double* x, double* y);
int dor_overland_conductance(dor_ctx* ctx, double* ke, double* kn, double* kw, double* ks);
/* a2d: 5*S values of the reference overland matrix (ring rows identity), rhs: S. */
int dor_assemble_overland(dor_ctx* ctx, double* a2d, double* rhs);
int dor_set_threads(int32_t nthreads);

/* ---- discrete adjoint (dhara_oracle_adjoint.cpp), same contracts as the dhara_ ones --- */
int dor_observe(dor_ctx* ctx, int32_t nsteps, double* q, double* w);
int dor_adjoint_run(dor_ctx* ctx, const dhara_loss_spec* spec, double* loss, double* grad_lnks,
                    double* grad_alpha, double* grad_n, dhara_step_stats* stats);
int dor_adjoint_residual(dor_ctx* ctx, const double* psi_new, double* F);
int dor_adjoint_jvp(dor_ctx* ctx, const double* psi_new, int32_t transpose, const double* v,
                    double* out);
/* dK/dpsi, dtheta/dalpha, dtheta/dn, dK/dalpha, dK/dn of each cell at psi (any may be NULL). */
int dor_vg_derivatives(dor_ctx* ctx, const double* psi, double* dK_dpsi, double* dth_da,
                       double* dth_dn, double* dK_da, double* dK_dn);

/* ---- CUSP-semantics building blocks (known-answer tests) ---------------------------- */
/* cusp::gallery::poisson5pt(A, m, n) / poisson7pt(A, m, n, k) as DIA
 * (cusp/gallery/detail/poisson.inl:28-86, stencil.inl:143-186). */
int dor_poisson5pt(int32_t m, int32_t n, int32_t* offsets, double* values);
int dor_poisson7pt(int32_t m, int32_t n, int32_t k, int32_t* offsets, double* values);
/* y = A x for a DIA matrix with pitch = nrows (dia_spmv.h:60-111). */
int dor_dia_spmv(int32_t nrows, int32_t ndiag, const int32_t* offsets, const double* values,
                 const double* x, double* y);
/* CUSP krylov drivers with monitor(b, maxit, rtol, atol); return iteration count in *iters
 * and the last residual norm in *rnorm; *converged follows monitor::converged(). */
int dor_dia_cg(int32_t nrows, int32_t ndiag, const int32_t* offsets, const double* values,
               const double* b, double* x, int32_t maxit, double rtol, double atol,
               int32_t* iters, double* rnorm, int32_t* converged);
int dor_dia_bicgstab(int32_t nrows, int32_t ndiag, const int32_t* offsets,
                     const double* values, const double* b, double* x, int32_t maxit,
                     double rtol, double atol, int32_t* iters, double* rnorm,
                     int32_t* converged);
int dor_dia_pcg_jacobi(int32_t nrows, int32_t ndiag, const int32_t* offsets,
                       const double* values, const double* b, double* x, int32_t maxit,
                       double rtol, int32_t* iters, double* rnorm, int32_t* converged);
/* Column preconditioner and PCG with it on a 7-point DIA matrix of an nx x ny x nz grid
 * (offsets -nx*ny, -nx, -1, 0, 1, nx, nx*ny; values 7 x n, pitch n). */
int dor_line_solve(int32_t nx, int32_t ny, int32_t nz, const double* values, const double* r, double* z);
int dor_dia_pcg_line(int32_t nx, int32_t ny, int32_t nz, const double* values, const double* b, double* x,
                     int32_t maxit, double rtol, int32_t* iters, double* rnorm, int32_t* converged);
/* Dense LU with partial pivoting, A row-major n*n (overwritten), b -> x. */
int dor_dense_solve(int32_t n, double* A, const double* b, double* x);
/* Thomas algorithm: a sub-, b main-, c super-diagonal, d rhs, x out (all length n). */
int dor_thomas(int32_t n, const double* a, const double* b, const double* c, const double* d,
               double* x);
/* Independent 1D column Richards solver (same discretisation, tridiagonal + Thomas):
 * nz cells of thickness dz, uniform Ks, top flux qtop [m/s] (+down), free drainage bottom.
 * psi: in initial profile, out after nsteps.  theta_out may be NULL. */
int dor_column_run(int32_t nz, double dz, const dhara_params* params, double Ks, double qtop,
                   int32_t nsteps, double* psi, double* theta_out, int32_t* picard_iters);

#ifdef __cplusplus
}
#endif
#endif
