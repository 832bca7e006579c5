/*
 * drd_oracle.h -- CPU oracle (test infrastructure only).
 *
 * A plain-C FP64 restatement of the reference's implicit-solve path,
 * used only as the checker by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  Nothing in the product links it.
 *
 * The oracle mirrors the C-ABI of include/drd_gpu.h with an "oc_" prefix
 * and host-resident state, so a test can drive both the same way.
 */
#ifndef DRD_ORACLE_H
#define DRD_ORACLE_H

#include <stdint.h>
#include "../include/drd_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oc_ctx oc_ctx;

void oc_params_default(drd_params* p, int32_t nx, int32_t ny);
int oc_init(const drd_params* prm, oc_ctx** out);
int oc_finalize(oc_ctx* c);
/* direct pointer to a module-`global` field (host memory owned by c)     */
double* oc_field(oc_ctx* c, int32_t field);
/* periodic x (lr_bd_type 'p'): the periodic ghost columns of a field written
 * through oc_field (no-op otherwise) */
int oc_fill_ghosts(oc_ctx* c, int32_t field);
int64_t oc_field_size(const drd_params* p, int32_t field);
/* eigenvalue tables of setup (TP/setup.f90:108-155)                      */
const double* oc_pre_lcc(oc_ctx* c); /* (0:nx,0:ny) */
/* ESOLP tables TP/setup.f90:157-235: 0..2 = ua,ub,uc (1:ny,0:nx),
 * 3..5 = va,vb,vc (1:ny-1,0:nx-1) */
const double* oc_esolp_table(oc_ctx* c, int32_t which);
const double* oc_ph_lcc(oc_ctx* c);  /* (0:nx-1,0:ny-1) */

int oc_velosolve_combined(oc_ctx* c, const double* psi, double* u_after,
                          double* v_after, const double* rmu_after,
                          const double* ph_after, double t,
                          const double* pf1, const double* pf2,
                          drd_solve_info* info);
int oc_velosolve_2_combined(oc_ctx* c, const double* psi1, const double* psi2,
                            const double* u_pre, double* u_after,
                            const double* v_pre, double* v_after,
                            const double* ph_pre, const double* ph_after,
                            const double* rmu_pre, const double* rmu_after,
                            double t, const double* pf1, const double* pf2,
                            drd_solve_info* info);
/* the momentum right-hand side only (rhs: Nuv vector), bdf 1 or 2        */
int oc_velo_rhs(oc_ctx* c, int32_t bdf, const double* psi1,
                const double* psi2, const double* u_pre, const double* v_pre,
                const double* ph_after, const double* rmu_after, double t,
                const double* pf1, const double* pf2, double* rhs);
int oc_presolve_2(oc_ctx* c, double* psi, double coef, double* mean_div);
int oc_ph_mu_solve(oc_ctx* c, double* ph_after, double* rmu_after, double t,
                   drd_solve_info* info);
/* ph_mu_solve right-hand side only (rhs: 2*nx*ny), also sets ph_M_0     */
int oc_ph_mu_rhs(oc_ctx* c, double* rhs);
int oc_particle_force(oc_ctx* c, const double* bx, const double* by, const double* br, const double* fx,
                      const double* fy, int32_t nbody, double lx, double ly);
int oc_particle_velocity(oc_ctx* c, const double* bx, const double* by, const double* br, int32_t nbody, double lx,
                         double ly, double* u_aver, double* v_aver);
int oc_interpolate_coef(oc_ctx* c, double lambda_eta, const double* px, const double* py, int32_t np,
                        double radius, const double* sx, const double* sy, int32_t nswm, double swm_radius);
int oc_phsolve(oc_ctx* c, double* rmu_after, double* ph_after, double t,
               drd_solve_info* info);
/* PhSolve_ori TP/PhSolve_ori.f90:1-234 (alpha_s: the wall stability term of
 * module global) and the CH re-initialisation of a solid domain
 * MF/finit.f90:132-188 = TP/reinint_by_PhSolve.f90 */
int oc_phsolve_ori(oc_ctx* c, double alpha_s, double* rmu_after, double* ph_after, drd_solve_info* info);
int oc_reinit_solid(oc_ctx* c, int32_t steps, double dt, double alpha_s, drd_solve_info* info);

/* y = alpha A x + beta y; with beta == 0, y is not read (as in BLAS), so it
 * may hold anything, NaN included                                          */
int oc_matvec(oc_ctx* c, int32_t op, double alpha, const double* x,
              double beta, double* y);
int oc_psolve(oc_ctx* c, int32_t prec, int32_t op, double* x_out,
              const double* b_in);
int oc_bicgstab(oc_ctx* c, int32_t op, int32_t prec, int32_t n,
                const double* b, double* x, int32_t* iter, double* resid,
                int32_t* info);
/* BiCGSTAB on a dense column-major n x n matrix with identity or Jacobi
 * preconditioner (prec 1 / 0): used to test the control flow alone.    */
int oc_bicgstab_dense(int32_t n, const double* a, int32_t jacobi,
                      const double* b, double* x, int32_t* iter,
                      double* resid, int32_t* info);
/* restarted flexible GMRES, right preconditioned (Saad 1993) */
int oc_fgmres(oc_ctx* c, int32_t op, int32_t prec, int32_t n, int32_t restart, const double* b, double* x,
              int32_t* iter, double* resid, int32_t* info);
int oc_fgmres_dense(int32_t n, const double* a, int32_t jacobi, int32_t restart, const double* b, double* x,
                    int32_t* iter, double* resid, int32_t* info);
int oc_step(oc_ctx* c, int32_t bdf, double t, drd_step_info* info);

/* transforms (definitions, O(n^2)) */
/* FFTW REDFT10 / REDFT01 of a 2-D nx x ny array (x fastest), in place  */
void oc_dct2_2d(int32_t nx, int32_t ny, double* a);
void oc_dct3_2d(int32_t nx, int32_t ny, double* a);
/* VCOST (LIB/vfftpack.f:643-704): m sequences of length n, stride 1
 * between elements of one sequence and `ld` between sequences          */
void oc_vcost(int32_t m, int32_t n, double* x, int32_t ld);
/* DCT(nr,nz,fO,fN,cof) TP/DCT.f90:1-57 ; fO,fN (-1:nr+1,-1:nz+1)        */
void oc_dct_solve(int32_t nr, int32_t nz, double* fO, double* fN,
                  const double* cof);
/* ESOLP_x TP/ESOLP.f90:1-82 ; pa,pb,pc (1:nz+1,0:nr); b,p (-1:nr+1,-1:nz+1) */
void oc_esolp_x(int32_t nr, int32_t nz, const double* pa, const double* pb,
                const double* pc, const double* b, double* p);
/* rfords TP/setup.f90:327-354 */
void oc_rfords(int32_t n, double* ff, const double* aa, const double* bb,
               const double* cc);
/* ch_bulk_energy TP/ch_bulk_energy.f90, MF/ch_bulk_energy.f90 and
 * ch_bdr_condition TP/ch_bdr_condition.f90                              */
/* canonical (pairwise, zero-padded to a power of two) sum of x[i]*y[i],
 * or of x[i] when y is NULL */
double oc_pairwise(int64_t n, const double* x, const double* y);
/* FFT-factored 2-D DCT of the pressure solve: mode 0 REDFT10 (x then y),
 * 1 REDFT01 (y then x) */
void oc_dct_2d_fast(int32_t nx, int32_t ny, double* a, int32_t mode);
/* FFTW R2HC (dir 0) / HC2R (dir 1, unnormalised) of one line, n a power of two */
void oc_rfft_line(int32_t n, double* x, int32_t dir);
double oc_ch_bulk_energy(double phi, double psi, int32_t mode);
double oc_ch_bulk_energy_mf(double phi, double psi);
double oc_ch_bdr_condition(double angle, double phi, int32_t mode);
/* op_double_gradient_Neum TP/matrix_tool.f90:178-219 on (-1:nx,-1:ny)   */
void oc_op_double_gradient_neum(int32_t nx, int32_t ny, double dx2,
                                double dy2, const double* rmu,
                                const double* mo, double* rhs);

#ifdef __cplusplus
}
#endif
/* OpenMP threads of the oracle's parallel loops (results do not depend on it) */
void oc_set_threads(int32_t n);
/* drd_set_initial_guess of the oracle: 0 previous step, 1 2x^n - x^{n-1} */
int oc_set_initial_guess(oc_ctx* c, int32_t mode);
int32_t oc_get_threads(void);

#endif
