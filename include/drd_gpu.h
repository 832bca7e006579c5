/*
 * drd_gpu.h -- C-ABI of the MI355X (gfx950) implicit-solve path of DRD-Micro.
 *
 * One process, one GPU.  Every entry point returns 0 on success and a
 * negative DRD_E* code on failure (drd_last_error() holds the message).
 * Arrays cross this boundary as plain pointers to contiguous FP64 data in
 * Fortran (column-major, x fastest) order with the reference's ghost layers:
 *
 *   cell field  f(-1:nx, -1:ny)    (nx+2)*(ny+2) doubles   TP/global.f90:56-59
 *   u field     u(-1:nx+1, -1:ny)  (nx+3)*(ny+2) doubles   TP/global.f90:54
 *   v field     v(-1:nx, 0:ny)     (nx+2)*(ny+1) doubles   TP/global.f90:54
 *   velocity Krylov vector  Nuv = (nx+1)*ny + nx*(ny-1)    TP/matrix_tool.f90:5-26
 *   phi/mu Krylov vector    2*nx*ny                        TP/matrix_tool.f90:82-98
 *   mu Krylov vector        nx*ny                          TP/matrix_tool.f90:142-155
 *
 * nx, ny are the reference's nr, nz (TP/global.f90:15-16), which the
 * reference fixes at compile time; here they are run-time fields of
 * drd_params.
 */
#ifndef DRD_GPU_H
#define DRD_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DRD_ABI_VERSION 1

/* ---- error codes ---------------------------------------------------- */
#define DRD_OK 0
#define DRD_E_ARG (-1)     /* bad argument or size                     */
#define DRD_E_HIP (-2)     /* HIP runtime error                        */
#define DRD_E_STATE (-3)   /* call out of order (e.g. ctx not set up)  */
#define DRD_E_UNSUPPORTED (-4)

/* ---- model variants -------------------------------------------------- */
/* TP: two_phase_flow_complex_domain; MF: microfluidics_complex_domain.     */
#define DRD_VARIANT_TP 0
#define DRD_VARIANT_MF 1

/* ---- preconditioner choices for the velocity solve -------------------- */
#define DRD_PREC_REFERENCE 0 /* what the reference uses: identity for BDF2,
                                Jacobi for BDF1 (TP), Jacobi (MF)         */
#define DRD_PREC_IDENTITY 1  /* TP/VeloSolve_2_combined.f90:779-780       */
#define DRD_PREC_JACOBI 2    /* TP/VeloSolve_combined.f90:679-741         */
#define DRD_PREC_DCT32 3     /* FP32 DCT/DST fast-Poisson preconditioner
                                with diagonal scaling (no reference
                                counterpart; power-of-two nx, ny)         */
#define DRD_PREC_DCT64 4     /* the same in FP64                          */
/* DRD_PREC_LEARNED (5): FP32 DCT + learned correction, defined below      */
/* DRD_PREC_MG (6), DRD_PREC_MG32 (7): geometric multigrid V-cycle, below */

/* ---- preconditioner choices for the Cahn-Hilliard phi/mu solve -------- */
#define DRD_CH_PREC_REFERENCE 0 /* scalar Jacobi TP/ph_mu_solve.f90:304-356 */
#define DRD_CH_PREC_BLOCK 1     /* exact 2x2 per-cell block Jacobi (no
                                   reference counterpart)                   */
#define DRD_CH_PREC_RBGS 2      /* sixteen red-black block SOR sweeps (weight
                                   1.25) of the same 2x2 blocks (no
                                   reference counterpart)                   */
#define DRD_CH_PREC_RBGS32 3    /* the same sweeps in FP32: mobilities, b,
                                   iterates and arithmetic (reciprocals
                                   where the FP64 form divides); FP64 b in,
                                   FP64 result out                          */

/* ---- outer Krylov method ------------------------------------------------ */
#define DRD_KRYLOV_BICGSTAB 0 /* LIB/BiCGSTAB.f, reference stopping rule    */
#define DRD_KRYLOV_FGMRES 1   /* restarted flexible GMRES (Saad 1993)       */

/* ---- run-time copy of the reference's module `global` scalars -------- */
typedef struct drd_params {
  int32_t nx, ny;          /* nr, nz                        TP/global.f90:15-16 */
  int32_t variant;         /* DRD_VARIANT_*                                      */
  int32_t phi_mode;        /* 1: phi in [0,1]; 2: phi in [-1,1] TP/mcl6.f90     */
  int32_t lr_bd_type;      /* 'n' (Neumann), 'd' (Dirichlet inflow) or 'p'
                            * (periodic x: no TP counterpart; packing as
                            * AM/matrix_tool.f90:81-139, pressure by a real
                            * FFT in x and REDFT10/01 in y as TP/CFT.f90:4-57;
                            * u has nx unknowns per row; every field keeps
                            * periodic ghost columns, which drd_set_field
                            * fills; TP variant, power-of-two nx, one
                            * instance, the Jacobi / identity / phi-mu block
                            * preconditioners)                                 */
  int32_t velo_prec;       /* DRD_PREC_*                                         */
  int32_t maxit_velo1;     /* 800   TP/VeloSolve_combined.f90:417                */
  int32_t maxit_velo2;     /* 2000  TP/VeloSolve_2_combined.f90:413              */
  int32_t maxit_phmu;      /* 5000  TP/ph_mu_solve.f90:156                       */
  int32_t maxit_ph;        /* 500   MF/PhSolve.f90:72                            */
  int32_t ch_prec;         /* DRD_CH_PREC_*                                      */
  int32_t krylov;          /* DRD_KRYLOV_* for the velocity solve                */
  double dx, dz;           /* dr, dz                                             */
  double dt, re, ca, xld;  /* time step, Reynolds, capillary, mobility           */
  double s_implicit;       /* stabilisation s                                   */
  double I_thickness;      /* interface thickness epsilon                       */
  double ggx, ggz;         /* body force                                         */
  double uwT, uwB;         /* wall speeds                                        */
  double xlsB;             /* slip length; must be < 1e-8 (no-slip branch)       */
  double tol;              /* Krylov tolerance                                   */
  double penalty_k;        /* diffuse-resistance coefficient TP/global.f90:108   */
  double angle;            /* static contact angle (degrees)                     */
  double lambda_rho;       /* density ratio; gamma1 = min(1, lambda_rho) TP/mcl6.f90 */
  double bdr_ph_used_adjust; /* 1e-6  TP/finit.f90:572                           */
  double twall, UU, bdr_upp, bdr_low; /* inflow profile ('d' only)              */
} drd_params;

/* fills p with the reference's TP defaults for an nx x ny grid            */
void drd_params_default(drd_params* p, int32_t nx, int32_t ny);

/* ---- per-solve report (INFO/ITER/RESID of LIB/BiCGSTAB.f:15-16) ------ */
typedef struct drd_solve_info {
  int32_t iter;      /* ITER out                                         */
  int32_t info;      /* INFO out: 0 ok, 1 maxit, -10/-11 breakdown        */
  double resid;      /* RESID out (recursive residual / ||b||)            */
  double true_resid; /* ||b - A x|| / ||b|| recomputed in FP64 (addition) */
} drd_solve_info;

typedef struct drd_step_info {
  drd_solve_info phmu;  /* ph_mu_solve                                   */
  drd_solve_info velo;  /* VeloSolve_(2_)combined                        */
  double div_before;    /* mean divergence printed by PreSolve_2:23       */
  double div_after_max; /* max |div u| after projection check (addition)  */
} drd_step_info;

typedef struct drd_ctx drd_ctx;

/* ---- field identifiers for module-`global` state ---------------------- */
enum drd_field {
  DRD_F_U = 0,          /* u(-1:nx+1,-1:ny)   */
  DRD_F_V = 1,          /* v(-1:nx,0:ny)      */
  DRD_F_P = 2,          /* cell fields ...    */
  DRD_F_PH = 3,
  DRD_F_RMU = 4,
  DRD_F_RHO = 5,
  DRD_F_ETA = 6,
  DRD_F_BDR_PH = 7,
  DRD_F_BDR_PH_USED = 8,
  DRD_F_BDR_PH_PENALTY = 9,
  DRD_F_BDR_PH_GRAD = 10,
  DRD_F_PH_M_INNER = 11,
  DRD_F_PH_M_0 = 12,
  DRD_F_ANGLEB = 13,    /* angleB(0:nx)       */
  DRD_F_ANGLET = 14,    /* angleT(0:nx)       */
  /* driver state kept on the device by drd_step (TP/mcl6.f90:21-30)     */
  DRD_F_U_PRE = 15,
  DRD_F_V_PRE = 16,
  DRD_F_PH_PRE = 17,
  DRD_F_RMU_PRE = 18,
  DRD_F_PSI1 = 19,
  DRD_F_PSI2 = 20,
  DRD_F_FPD_PH = 21,    /* fpd_ph(-1:nx,-1:ny): solid and particle indicator
                           (drd_interpolate_coef)                        */
  DRD_F_PFORCE1 = 22,   /* particle_force1, particle_force2 (-1:nx,-1:ny): the  */
  DRD_F_PFORCE2 = 23,   /* body force drd_step adds to the momentum equation
                           (TP/VeloSolve_2_combined.f90:242); drd_particle_force */
  DRD_F_COUNT = 24
};

/* number of doubles of a field for params p (0 for an unknown id)        */
int64_t drd_field_size(const drd_params* p, int32_t field);

/* ---- lifecycle --------------------------------------------------------
 * drd_init replaces PreSolve_init (TP/PreSolve_init.f90:1-16, FFTW plans)
 * and setup (TP/setup.f90:1-267, eigenvalue tables): it creates the device
 * context, DCT twiddles and Pre_Lcc / Ph_Lcc tables for prm.           */
int drd_init(const drd_params* prm, drd_ctx** out);
int drd_finalize(drd_ctx* ctx);
/* replaces the scalars of an existing context (same nx, ny); rebuilds the
 * eigenvalue tables when dt, dx, dz or the CH coefficients change          */
int drd_set_params(drd_ctx* ctx, const drd_params* prm);
const char* drd_last_error(void);
int drd_abi_version(void);

/* copy one module-`global` field host <-> device                         */
int drd_set_field(drd_ctx* ctx, int32_t field, const double* host);
int drd_get_field(drd_ctx* ctx, int32_t field, double* host);

/* ---- the Fortran subroutines, same dummy arguments ---------------------
 * Hidden inputs come from the context (drd_set_field); outputs are host
 * arrays.  Each call uploads its array arguments, runs on the GPU and
 * downloads its outputs before returning (synchronous, like the Fortran). */

/* TP/VeloSolve_combined.f90:5-18 (BDF1 momentum predictor)               */
int drd_velosolve_combined(drd_ctx* ctx, const double* psi, double* u_after,
                           double* v_after, const double* rmu_after,
                           const double* ph_after, double t,
                           const double* particle_force1,
                           const double* particle_force2,
                           drd_solve_info* info);

/* TP/VeloSolve_2_combined.f90:4-20 (BDF2 momentum predictor)             */
int drd_velosolve_2_combined(drd_ctx* ctx, const double* psi1,
                             const double* psi2, const double* u_pre,
                             double* u_after, const double* v_pre,
                             double* v_after, const double* ph_pre,
                             const double* ph_after, const double* rmu_pre,
                             const double* rmu_after, double t,
                             const double* particle_force1,
                             const double* particle_force2,
                             drd_solve_info* info);

/* TP/PreSolve_2.f90:1-130: psi from div(u,v) of the context's u, v.      */
int drd_presolve_2(drd_ctx* ctx, double* psi, double coef,
                   double* mean_div);

/* TP/ph_mu_solve.f90:3-197 (coupled phi/mu Cahn-Hilliard)                */
int drd_ph_mu_solve(drd_ctx* ctx, double* ph_after, double* rmu_after,
                    double t, drd_solve_info* info);

/* MF/PhSolve.f90:1-126 (mu-only Cahn-Hilliard, DCT-I preconditioner)    */
int drd_phsolve(drd_ctx* ctx, double* rmu_after, double* ph_after, double t,
                drd_solve_info* info);

/* ---- the MATVEC / PSOLVE callbacks (y := alpha*A*x + beta*y) ---------- */
#define DRD_OP_VELO1 0 /* MatrixVector_uv_CG   TP/VeloSolve_combined.f90:537-675   */
#define DRD_OP_VELO2 1 /* MatrixVector_uv_2_CG TP/VeloSolve_2_combined.f90:536-676 */
#define DRD_OP_PHMU 2  /* MATRIX_VECTOR_PH_MU_CG TP/ph_mu_solve.f90:203-299        */
#define DRD_OP_PH 3    /* MATRIX_VECTOR_CG     MF/PhSolve.f90:240-292              */
#define DRD_OP_PHORI 4 /* MATRIX_VECTOR_CG_ori TP/PhSolve_ori.f90:348-397 (= MF)   */
int drd_matvec(drd_ctx* ctx, int32_t op, double alpha, const double* x,
               double beta, double* y);

#define DRD_PS_VELO1_JACOBI 0 /* PSolve_uv_CG   TP/VeloSolve_combined.f90:679-741   */
#define DRD_PS_IDENTITY 1     /* PSolve_uv_2_CG TP/VeloSolve_2_combined.f90:779-780 */
#define DRD_PS_PHMU_JACOBI 2  /* PSolve_PH_MU_CG TP/ph_mu_solve.f90:304-356         */
#define DRD_PS_PH_DCT 3       /* PSolve_CG      MF/PhSolve.f90:296-313              */
#define DRD_PS_VELO_DCT32 4   /* FP32 DCT/DST velocity preconditioner (addition)    */
#define DRD_PS_PHMU_BLOCK 5   /* 2x2 block Jacobi for the phi/mu system (addition)   */
#define DRD_PS_VELO_DCT64 6   /* FP64 DCT/DST velocity preconditioner (addition)    */
#define DRD_PS_VELO_ESOLP 7   /* PSolve_uv_CG PD/VeloSolve_combined.f90:697-807
                                 (ESOLP_x TP/ESOLP.f90 on both blocks)            */
/* DRD_PS_VELO_LEARNED (8): FP32 DCT + learned correction, defined below   */
/* DRD_PS_VELO_MG (9), DRD_PS_VELO_MG32 (11): multigrid V-cycle, below     */
#define DRD_PS_PHMU_RBGS 10   /* red-black block Gauss-Seidel for the phi/mu system
                                 (addition): x = 0; then 8 times: the red cells,
                                 then the black cells, x_c += B_c^-1 (b - A x)_c
                                 with B_c the exact 2x2 block of cell c         */
#define DRD_PS_PHMU_RBGS32 12 /* DRD_PS_PHMU_RBGS in FP32 (bu, M_inner, M_0,
                                 b, iterates, arithmetic; reciprocals for
                                 divisions); FP64 in and out (addition)      */
int drd_psolve(drd_ctx* ctx, int32_t prec, int32_t op, double* x_out,
               const double* b_in);

/* PhSolve_ori(rmu_after, ph_after, t) TP/PhSolve_ori.f90:1-234 (the same in
 * MF/): one implicit CH step for phi with unit mobility (its operator is
 * DRD_OP_PHORI, preconditioned by DRD_PS_PH_DCT) from the context's ph, u,
 * v and contact angles; alpha_s is the wall term of module global
 * (MF/finit.f90:374), wall types 'n', Neumann sides, at most 500
 * iterations (:148).  Outputs are host cell fields.                        */
int drd_phsolve_ori(drd_ctx* ctx, double alpha_s, double* rmu_after, double* ph_after,
                    drd_solve_info* info);

/* The CH re-initialisation of a solid domain, MF/finit.f90:132-188 (=
 * TP/reinint_by_PhSolve.f90:1-91), on the context's fields: ph := bdr_ph
 * (ghosts copied), ph_M_0 = ph_M_inner = 1, `steps` PhSolve_ori steps of
 * size dt (the reference: 51 steps of 1e-4), then bdr_ph = (ph + 1) / 2
 * (the interior clamped to [0, 1]) and bdr_ph_penalty = [bdr_ph > 0.5].
 * ph and rmu keep the last step's state and the mobilities stay 1, as the
 * reference leaves them; the context's dt (and its Ph_Lcc) is unchanged
 * afterwards.  info: the last step's solve.  Single instance, Neumann
 * sides (the reference forces lr_bd_type 'n' for it).                    */
int drd_reinit_solid(drd_ctx* ctx, int32_t steps, double dt, double alpha_s, drd_solve_info* info);

/* BICGSTAB(N,B,X,WORK,LDW,ITER,RESID,MATVEC,PSOLVE,INFO) LIB/BiCGSTAB.f:2-3
 * with the MATVEC/PSOLVE callbacks chosen by id.  *iter in: max
 * iterations, out: iterations; *resid in: tol, out: measure.             */
int drd_bicgstab(drd_ctx* ctx, int32_t op, int32_t prec, int32_t n,
                 const double* b, double* x, int32_t* iter, double* resid,
                 int32_t* info);

/* BICGSTAB(N,B,X,WORK,LDW,ITER,RESID,MATVEC,PSOLVE,INFO) LIB/BiCGSTAB.f:2-22
 * with the caller's own operator and preconditioner: MATVEC(ALPHA, X, BETA,
 * Y) computes Y = ALPHA*A*X + BETA*Y and PSOLVE(X, B) X = M^-1 B, on host
 * vectors, every argument by reference as a Fortran caller passes them.
 * The Krylov vectors, updates, dot products and stopping rule run on the
 * device (the state machine of drd_bicgstab, canonical reductions); each
 * product and preconditioner application copies its operand to the host,
 * calls the routine and copies the result back.  MATVEC is called as
 * LIB/BiCGSTABREVCOM.f calls it (r0 with ALPHA = -1, BETA = 1 on b; the
 * iteration's products with BETA = 0 into a zeroed Y), and neither routine
 * is called after convergence or an early exit.  psolve NULL: identity.
 * INFO -1, -2, -3 for N < 0, LDW < max(1, N), ITER <= 0, as the reference
 * (nothing is called); N = 0 returns ITER = 0, RESID = 0.  ITER in: MAXIT,
 * RESID in: the tolerance.  ctx NULL: the library's own context, sized for
 * N (up to 2 * 4096^2); calls on it are serialised.                       */
typedef void (*drd_matvec_cb)(const double* alpha, const double* x, const double* beta, double* y);
typedef void (*drd_psolve_cb)(double* x, const double* b);
int drd_bicgstab_cb(drd_ctx* ctx, int32_t n, const double* b, double* x, int32_t ldw, int32_t* iter,
                    double* resid, drd_matvec_cb matvec, drd_psolve_cb psolve, int32_t* info);

/* FGMRES(m) with right preconditioning (Saad 1993); no reference code
 * (LIB/fgmres.f wraps a closed library).  Same argument meaning as
 * drd_bicgstab; restart = m.                                               */
int drd_fgmres(drd_ctx* ctx, int32_t op, int32_t prec, int32_t n,
               int32_t restart, const double* b, double* x, int32_t* iter,
               double* resid, int32_t* info);

/* ---- device-resident time stepping (TP/mcl6.f90:395-696 sequence) ----
 * bdf=1: first step (ph_mu_solve, VeloSolve_combined, PreSolve_2 with
 * coef re*gamma/dt); bdf=2: later steps.  State lives on the device
 * (drd_set_field/drd_get_field move it).                                   */
int drd_step(drd_ctx* ctx, int32_t bdf, double t, drd_step_info* info);
int drd_synchronize(drd_ctx* ctx);

/* ---- snapshots in the reference's data-file format ----------------------
 * drd_write_snapshot replaces output (TP/inout.f90:1-20): Fortran
 * unformatted sequential records (4-byte markers) x, y, u, v, p, f, f2, t,
 * each array written as ((a(i,k), k=0,ny), i=0,nx), with x, y the cell
 * centres; f and f2 are cell fields chosen by id (the reference writes
 * fpd_ph and ph, or bdr_ph and ph).  TP/post_process/da1.py reads it.
 * drd_read_snapshot replaces input (TP/inout.f90:41-61) but reads all seven
 * arrays (the reference's input skips f2) and sets u, v, p, f, f2 at their
 * (0:nx, 0:ny) elements; *t receives the time.                              */
int drd_write_snapshot(drd_ctx* ctx, const char* path, int32_t f_field, int32_t f2_field, double t);
int drd_read_snapshot(drd_ctx* ctx, const char* path, int32_t f_field, int32_t f2_field, double* t);

/* ---- batches of independent instances (no reference counterpart) --------
 * One drd_batch holds `count` independent instances (each its own grid,
 * fields and solver state).  Instances of equal nx x ny form a group that
 * steps together: every kernel of the step is one launch over all of the
 * group's instances (the instance is the launch's z index), each instance
 * keeping its own BiCGSTAB state machine, so finished instances idle while
 * the others iterate.  Instances of one size must share every drd_params
 * field; per-instance resistance goes into each instance's bdr_ph_penalty
 * field (the operator uses penalty_k * bdr_ph_penalty).  BiCGSTAB only;
 * velocity preconditioner DCT32, DCT64 (power-of-two grids), MG, MG32 (even
 * grids of at least 128 x 64: each instance's hierarchies sit in its own
 * block of the group's memory), Jacobi, identity or reference.             */
typedef struct drd_batch drd_batch;
int drd_batch_create(int32_t count, const drd_params* params, drd_batch** out);
int drd_batch_destroy(drd_batch* b);
/* one instance's field (layout as drd_set_field)                          */
int drd_batch_set_field(drd_batch* b, int32_t inst, int32_t field, const double* host);
int drd_batch_get_field(drd_batch* b, int32_t inst, int32_t field, double* host);
/* drd_step for every instance; infos[count] receives each one's report     */
int drd_batch_step(drd_batch* b, int32_t bdf, double t, drd_step_info* infos);
/* number of groups; desc[3g..3g+2] = nx, ny, instances of group g (g < cap) */
int drd_batch_groups(drd_batch* b, int32_t* ngroups, int32_t* desc, int32_t cap);

/* ---- captured iteration chunks (no reference counterpart) ----------------
 * BiCGSTAB queues its iterations in chunks between polls of the state
 * machine (the first chunk one iteration shorter than the operator's
 * previous solve, then 1, 2, 4, ... 32).  With graphs on, each chunk of two
 * or more is a hipGraph captured on first use and replayed afterwards (one
 * graph per operator, preconditioner, chunk length and vector pair; kernels
 * only; replayed only while the pointers and scalars it baked in are
 * unchanged).  Results are bit for bit those without graphs.  Off by
 * default for contexts and batches.                                         */
int drd_use_graphs(drd_ctx* ctx, int32_t on);
/* Captured chunks held now (*live) and replays refused since the context was
 * made because something a chunk bakes in had changed (*recaptured; such a
 * chunk is captured again instead).                                          */
int drd_graph_count(drd_ctx* ctx, int32_t* live, int64_t* recaptured);
int drd_batch_use_graphs(drd_batch* b, int32_t on);

/* ---- multigrid V-cycle form (no reference counterpart) -------------------
 * on (the default): each leg of each level of the V(2,2) cycle of
 * DRD_PREC_MG / DRD_PREC_MG32 is one LDS-tiled launch (two sweeps, residual
 * and restriction going down; prolongation and two sweeps going up).  off:
 * one launch per sweep and transfer.  Both give the same values bit for bit;
 * the switch exists to show that, and to time the two forms.                */
int drd_mg_fused(drd_ctx* ctx, int32_t on);

/* ---- tile autotuner (no reference counterpart) ----------------------------
 * Times the DCT/DST velocity preconditioner with each transpose tile shape
 * (0: 32x32, 1: 16x64, 2: 64x16 elements) on the current fields and keeps
 * the fastest (results do not depend on the shape).  *shape = -1 when the
 * context has no DCT preconditioner; ms[3] receives the mean time per
 * application of each shape.  The batch form tunes every size group
 * (shapes[g], ms[3g..3g+2] for g < cap).                                     */
int drd_tune(drd_ctx* ctx, int32_t* shape, double* ms);
int drd_batch_tune(drd_batch* b, int32_t* shapes, double* ms, int32_t cap);

/* ---- initial guess of the BDF2 solves (no reference counterpart) ----------
 * mode 0 (default, the reference's): a solve starts from the previous step's
 * solution.  mode 1: from the extrapolation 2 x^n - x^{n-1} -- the velocity
 * solves of drd_step and drd_velosolve_2_combined (x^{n-1} = u_pre, v_pre)
 * and the phi/mu solve of drd_step (x^{n-1} = ph_pre, rmu_pre).  The stopping
 * test is unchanged; a smoothly varying solution needs fewer iterations.    */
int drd_set_initial_guess(drd_ctx* ctx, int32_t mode);
int drd_batch_set_initial_guess(drd_batch* b, int32_t mode);

/* ---- measurement hook (no reference counterpart) -----------------------
 * Launches one kernel of the solve path `reps` times on the context's
 * stream between HIP events and returns the mean launch duration and the
 * kernel's algorithmic HBM bytes per launch (every input read once, every
 * output written once).                                                    */
#define DRD_K_VELO_MATVEC 0 /* fused v = A p, (rtld, v); BDF2 operator     */
#define DRD_K_PHMU_MATVEC 1 /* fused CH block matvec + dot                 */
#define DRD_K_BICG_XR 2     /* x,r update + two dots                       */
#define DRD_K_PRESSURE 3    /* whole pressure projection                   */
#define DRD_K_COPY 4        /* plain copy of Nuv doubles (counter scale)    */
#define DRD_K_UNET_FWD 5    /* forward pass of the learned correction's
                               ny x nx network (after drd_learn_init);
                               *bytes returns its MFMA FLOPs instead      */
#define DRD_K_VELO_PREC 6   /* one application of the DCT/DST velocity
                               preconditioner (all its passes)             */
#define DRD_K_VELO_MG 7     /* one multigrid V-cycle (all its launches);
                               *bytes: FP64 in and out of the fine level   */
#define DRD_K_PHMU_SWEEP32 8 /* one middle sweep of DRD_PS_PHMU_RBGS32 (FP32
                                iterate in and out, FP32 b, bu, M_inner,
                                M_0), after one application of it         */
int drd_bench_kernel(drd_ctx* ctx, int32_t which, int32_t reps, double* avg_ms,
                     double* bytes);
/* Times kernel `which` inside the caller's own solves, in groups of
 * consecutive launches, each group bracketed by a pair of HIP events on the
 * context's stream (up to `capacity` groups; 0 disables and resets):
 *   DRD_K_VELO_MATVEC: the BDF2 operator's fused matvec with one dot, as
 *     BiCGSTAB launches it; a group is one launch, timed by its own
 *     dispatch timestamps (hipExtLaunchKernelGGL);
 *   DRD_K_PHMU_SWEEP32: the middle sweeps of DRD_PS_PHMU_RBGS32; a group is
 *     the six of one application, between two event markers on the stream.
 * drd_kernel_timer_read waits for the stream and returns the number of timed
 * launches, their summed duration and the kernel's algorithmic bytes per
 * launch.                                                                 */
int drd_kernel_timer(drd_ctx* ctx, int32_t which, int32_t capacity);

/* ---- coefficient update ----------------------------------------------------
 * interpolate_coef TP/interpolate_coef.f90:4-115 with dist TP/dist.f90:3-76
 * and fpd_dist TP/particles_move.f90:361-378, on the device: for every cell,
 *   fpd_ph = (tanh(-d / sqrt(2) / I_thickness) + 1) / 2 + bdr_ph,
 *   d = min(1e10, min_i |x - p_i| - radius, min_j |x - s_j| - swm_radius)
 * at the cell centres ((i + 0.5) dx, (k + 0.5) dz), with ghosts copied from
 * the nearest interior cell, then rho = 1 + (lambda_rho - 1) fpd_ph and
 * eta = 1 + (lambda_eta - 1) fpd_ph.  lambda_rho and I_thickness are the
 * context's; lambda_eta is not a drd_params field and comes as an argument.
 * px, py: the np particle centres (points(:,0), points(:,1)); sx, sy: the
 * nswm swimmer centres (pos_swm); host arrays, copied per call (2 (np +
 * nswm) at most (nx + 2)(ny + 2)).  Writes DRD_F_FPD_PH, DRD_F_RHO and
 * DRD_F_ETA; single-instance contexts only.                                 */
int drd_interpolate_coef(drd_ctx* ctx, double lambda_eta, const double* px, const double* py, int32_t np,
                         double radius, const double* sx, const double* sy, int32_t nswm, double swm_radius);

/* ---- particle coupling reductions -------------------------------------------
 * A body is a particle, a swimmer or a swimmer's phantom: centre (bx, by),
 * radius br, force (fx, fy).  Its weight at the cell centre ((i + 0.5) dx,
 * (k + 0.5) dz) is w = (tanh(-d / sqrt(2) / I_thickness) + 1) / 2, d the
 * distance in the periodic lx x ly box of fpd_dist_periodic
 * (TP/particles_move.f90:326-357).  Host arrays of nbody entries, copied per
 * call; 5 nbody at most max(nu + nv, (nx + 2)(ny + 2)); single-instance
 * contexts only.
 *
 * force_caculation_vel TP/particles_move.f90:758-839 after its call of
 * force_caculation (the per-body forces come in fx, fy): sets DRD_F_PFORCE1/2
 * to the sum over bodies, in order, of F w / (sum of w over the cells * dx dz),
 * zero in the ghosts.  The reference's three loops are one list: particles
 * (points, tot_F, fpd_radius), then swimmers (pos_swm, tot_F_swm, swm_radius),
 * then, with phantoms on, the phantoms (pos_swm - swm_bp_dist (cos, sin)
 * (active_F_swm_ang), -active_F_swm, swm_radius).                           */
int drd_particle_force(drd_ctx* ctx, const double* bx, const double* by, const double* br, const double* fx,
                       const double* fy, int32_t nbody, double lx, double ly);
/* update_position_vel TP/particles_move.f90:842-920, the grid part: the
 * w-weighted averages over the cells of (u(i,k) + u(i,k+1)) / 2 and
 * (v(i,k) + v(i,k+1)) / 2 from the context's u, v, into u_aver[nbody],
 * v_aver[nbody].  The position update and the periodic wrap stay with the
 * caller, as the particle ODE does.                                          */
int drd_particle_velocity(drd_ctx* ctx, const double* bx, const double* by, const double* br, int32_t nbody,
                          double lx, double ly, double* u_aver, double* v_aver);
/* No counterpart in the reference.  drd_particle_velocity for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc): no context; u (u field layout),
 * v (v field layout), bx, by, br and the outputs u_aver, v_aver are DEVICE
 * arrays; the grid comes from prm (nx, ny, dx, dz, I_thickness).  The work
 * buffers live in one fixed array of drd_noalloc_scratch_bytes() (16 MiB)
 * per GPU, whose address is looked up on the first call on each device;
 * nothing is allocated.  Rejected (DRD_E_ARG) beyond 8 nbody + 3 nbody *
 * chunks doubles of that array (chunks = min(ny, 256) row chunks, as
 * drd_particle_velocity takes them; e.g. nbody <= 2048 at ny >= 256).
 * The work is queued on `stream` (a hipStream_t; null: the null stream) and
 * the call returns without waiting.  Results equal drd_particle_velocity's
 * bit for bit.  Every call on one GPU shares that GPU's array, so calls that
 * overlap on one GPU (other streams, other host threads) race.            */
int64_t drd_noalloc_scratch_bytes(void);
int drd_particle_velocity_noalloc(const drd_params* prm, const double* u, const double* v, const double* bx,
                                  const double* by, const double* br, int32_t nbody, double lx, double ly,
                                  double* u_aver, double* v_aver, void* stream);
int drd_kernel_timer_read(drd_ctx* ctx, int32_t* count, double* total_ms, double* bytes);
/* The same groups one by one: ms[i] is the mean launch duration of timed
 * group i (i < min(*count, capacity)), in launch order; *count is the number
 * of groups.  A launch issued after its solve finished (the tail of a queued
 * chunk) returns at once and moves no data; its duration tells it apart.  */
int drd_kernel_timer_launches(drd_ctx* ctx, int32_t capacity, float* ms, int32_t* count);

/* ---- learned correction of the velocity preconditioner ------------------
 * No reference counterpart (an addition beside the drop-in).  DRD_PS_VELO_LEARNED
 * is the FP32 DCT/DST solve z = M r followed by the error correction of a
 * 3-level U-Net (16/32/32 channels, BF16 MFMA convolutions) on the ny x nx
 * face image:
 *   z' = z + alpha s_d N(r/s_r, d/s_d, z/s_z, log eta, chi),  d = r - A z,
 * s_* the RMS of each vector, N trained towards (A^-1 r - z)/s_d, and alpha
 * the step that minimises |r - A z'| (so |r - A z'| <= |d|).  The correction
 * is not linear in r, so only the flexible Krylov accepts it (drd_bicgstab
 * rejects it); drd_fgmres applies it at the first direction of each restart
 * cycle (or every drd_learn_set_every-th) and the plain FP32 DCT solve for
 * the cycle's other directions.
 * velo_prec = DRD_PREC_LEARNED makes drd_step solve the BDF2 velocity system
 * with it, judged once: the first solve after drd_learn_enable runs FGMRES
 * with the correction, FGMRES with the plain preconditioner and the plain
 * method (BiCGSTAB with the FP32 DCT preconditioner) from the same start, and
 * keeps the correction only if it needed no more iterations than plain FGMRES
 * and no more time than the plain method.  Otherwise, and whenever a later
 * learned solve needs more than 1.25x the judged plain FGMRES iterations, the
 * context falls back to the plain method (DRD_LEARN_FALLBACK).              */
#define DRD_PREC_LEARNED 5
#define DRD_PS_VELO_LEARNED 8

/* ---- multigrid velocity preconditioner -----------------------------------
 * No reference counterpart (an addition beside the drop-in).
 * DRD_PS_VELO_MG is one V-cycle of geometric multigrid on the BDF2 velocity
 * operator (even nx >= 128, ny >= 64):
 *   - the levels halve the grid while the sizes are even and the halves keep
 *     nx >= 64 and ny >= 32;
 *   - each coarse operator is the same MAC stencil with doubled spacings, on
 *     rho, eta, chi averaged over 2x2 cells;
 *   - the transfers are staggered bilinear prolongation P and R = P^T / 4;
 *   - Jacobi smoothing (exact diagonal; weight 0.7 on the finest level,
 *     1.2 on the coarser ones): two sweeps
 *     before the coarse correction and two after;
 *   - the FP64 DCT/DST fast-Poisson solve on the coarsest level: the FFT
 *     form when its sizes are powers of two, dense line transforms
 *     otherwise (a coarsest level of at most 256 x 256 cells).
 * It is a fixed linear map, so BiCGSTAB accepts it.  Its BiCGSTAB iteration
 * counts stay nearly constant as the grid is refined, unlike the fast-Poisson
 * preconditioner's.  velo_prec = DRD_PREC_MG makes drd_step solve the BDF2
 * velocity system with it (the first, BDF1 step takes its BDF1 form); batches
 * take it too.
 * DRD_PS_VELO_MG32 / DRD_PREC_MG32: the same cycle with every level but the
 * coarsest stored in FP32 (coefficients, diagonals, vectors; level 0 holds an
 * FP32 copy of rho, eta, chi), which halves the bytes of each sweep, and its
 * stencils computed in FP32 (reciprocals where the FP64 form divides).  b,
 * z, each sweep's update and the coarsest level stay FP64.                   */
#define DRD_PREC_MG 6
#define DRD_PS_VELO_MG 9
#define DRD_PREC_MG32 7
#define DRD_PS_VELO_MG32 11
/* The same fast-path preconditioners built for the BDF1 operator
 * MatrixVector_uv_CG (TP/VeloSolve_combined.f90:537-675: mass coefficient
 * re rho instead of 1.5 re rho, no penalty term); op must be DRD_OP_VELO1.
 * drd_step's first (BDF1) step takes them in the DCT, learned and multigrid
 * modes, where the reference's Jacobi needs more than its 800 iterations on
 * large grids.  No reference counterpart.                                    */
#define DRD_PS_VELO2_JACOBI 17 /* PSolve_uv_2_CG of MF MF/VeloSolve_2_combined.f90:662-711:
                                  Jacobi with mass coefficient 1.5 re, no penalty
                                  (the MF variant's BDF2 preconditioner)       */
#define DRD_PS_VELO1_DCT32 13
#define DRD_PS_VELO1_DCT64 14
#define DRD_PS_VELO1_MG 15
#define DRD_PS_VELO1_MG32 16

#define DRD_LEARN_OFF 0      /* not enabled: the plain method               */
#define DRD_LEARN_JUDGE 1    /* the next velocity solve is judged           */
#define DRD_LEARN_ON 2       /* correction in use                           */
#define DRD_LEARN_FALLBACK 3 /* regressed: the plain method                 */

typedef struct drd_learn_info {
  int32_t state;      /* DRD_LEARN_*                                        */
  int32_t pairs;      /* residual/error pairs held                          */
  int32_t capacity;   /* pairs the store holds (oldest replaced first)       */
  int32_t train_iters;/* Adam steps taken                                   */
  int32_t ref_iters;  /* plain-DCT FGMRES iterations of the judged solve    */
  int32_t judged_iters; /* learned FGMRES iterations of the judged solve    */
  int32_t last_iters; /* iterations of the last learned velocity solve      */
  int32_t fallbacks;  /* regressions seen                                   */
  double loss_first;  /* mean crop loss of the first training batch         */
  double loss_last;   /* mean crop loss of the last training batch          */
  double train_ms;    /* wall time of the last drd_learn_train call         */
  double judged_ms;   /* wall time of the judged learned FGMRES solve       */
  double plain_ms;    /* wall time of the judged plain-method solve         */
} drd_learn_info;

/* Allocates the pair store (capacity pairs of ny x nx images: 24 bytes per
 * cell each) and two networks sharing one parameter vector: an ny x nx one
 * for the preconditioner and a crop_h x crop_w one for training (multiples
 * of 4, at most ny x nx).  He-normal initial weights from seed.            */
int drd_learn_init(drd_ctx* ctx, int32_t capacity, int32_t crop_h, int32_t crop_w, uint64_t seed);
/* The same with a store of crops: each recorded iterate adds
 * crops_per_iterate pairs, random crop_h x crop_w crops of its images
 * (24 bytes per crop cell each), so the store holds many more pairs (20k
 * crops of 256 x 256 take 31 GB).  crops_per_iterate 0 is drd_learn_init.   */
int drd_learn_init_crops(drd_ctx* ctx, int32_t capacity, int32_t crop_h, int32_t crop_w,
                         int32_t crops_per_iterate, uint64_t seed);
/* FGMRES takes the correction at the directions j of a restart cycle with
 * j % every == 0 (every 1: every direction); 0 (the default): j == 0 only,
 * the cycle's other directions from the plain FP32 DCT solve.               */
int drd_learn_set_every(drd_ctx* ctx, int32_t every);
/* Solver-emitted pairs: solves A x = b (the BDF2 velocity operator of the
 * context's fields) from x = 0 by BiCGSTAB with the FP32 DCT preconditioner
 * to `tol`, keeping iterates x_k at k = 1, 2, 3, 5, 8, 12, ... (up to 0.8x
 * the final count).  Each gives r = b - A x_k, z = M r, d = r - A z,
 * e = x - x_k and the pair (features, (e - z)/s_d).  x (host, may be null)
 * receives the solution.                                                   */
int drd_learn_collect(drd_ctx* ctx, const double* b, double tol, double* x, int32_t* pairs_added,
                      int32_t* iters);
/* While on, drd_step's BDF2 velocity solves run as recorded BiCGSTAB with the
 * FP32 DCT preconditioner and add their pairs.                              */
int drd_learn_record(drd_ctx* ctx, int32_t on);
/* iters Adam steps (beta 0.9/0.999) over batches of `batch` random crops of
 * the stored pairs, learning rate lr decaying to lr/10 (cosine); the
 * preconditioner's network then takes the new parameters.                  */
int drd_learn_train(drd_ctx* ctx, int32_t iters, int32_t batch, double lr);
/* on: judge the next velocity solve (DRD_LEARN_JUDGE); off: DRD_LEARN_OFF   */
int drd_learn_enable(drd_ctx* ctx, int32_t on);
int drd_learn_get_info(drd_ctx* ctx, drd_learn_info* out);
/* copy the network parameters out (get) or in (set); n = drd_unet_param_count() */
int drd_learn_get_params(drd_ctx* ctx, float* params, int32_t n);
int drd_learn_set_params(drd_ctx* ctx, const float* params, int32_t n);

/* Stand-alone forward/backward of the network on host data (test hook):
 * in: H*W*8 BF16 bit patterns (NHWC), target: H*W*2 FP32 (may be null: no
 * backward), out: H*W*2, grads: drd_unet_param_count() floats of the loss
 * 0.5*sum((out-target)^2)/(2HW), *loss.  params_after (may be null) gets the
 * parameters after one Adam step (lr, beta 0.9/0.999, eps 1e-8).             */
int drd_unet_param_count(void);
int drd_unet_run(int32_t H, int32_t W, const float* params, const uint16_t* in, const float* target,
                 float* out, float* grads, double* loss, float lr, float* params_after);

#ifdef __cplusplus
}
#endif
#endif /* DRD_GPU_H */
