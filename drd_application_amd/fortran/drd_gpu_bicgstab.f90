! drd_gpu_bicgstab.f90 -- Fortran drop-in for the Krylov routine itself:
! BICGSTAB(N, B, X, WORK, LDW, ITER, RESID, MATVEC, PSOLVE, INFO) of
! LIB/BiCGSTAB.f:2-22 with the caller's own MATVEC and PSOLVE, over
! drd_bicgstab_cb of include/drd_gpu.h (the iteration on the GPU, the
! caller's routines called on host copies of their operands).
!
! Link it in place of LIB/BiCGSTAB.f and LIB/BiCGSTABREVCOM.f (and MKL's
! BLAS-1): every solve routine of the reference that calls BICGSTAB with its
! own matrix-vector product and preconditioner (TP/VeloSolve_combined.f90,
! TP/ph_mu_solve.f90, MF/PhSolve.f90, PD/diffusion_eqn.f90, ...) then runs
! its Krylov loop on the GPU unchanged.  It needs no module `global`; build
! it with drd_gpu_c.f90 and the caller's real kind.  WORK is not used (the
! workspace lives on the device); LDW is checked as the reference checks it.

! The caller's routines, called through C-interoperable trampolines.
module drd_bicgstab_cb_m
  use iso_c_binding
  implicit none
  procedure(), pointer, save :: drd_user_matvec => null()
  procedure(), pointer, save :: drd_user_psolve => null()
contains
  subroutine drd_cb_matvec(alpha, x, beta, y) bind(C)
    real(c_double), intent(in) :: alpha, beta
    real(c_double) :: x(*), y(*)
    call drd_user_matvec(alpha, x, beta, y)
  end subroutine drd_cb_matvec

  subroutine drd_cb_psolve(x, b) bind(C)
    real(c_double) :: x(*), b(*)
    call drd_user_psolve(x, b)
  end subroutine drd_cb_psolve
end module drd_bicgstab_cb_m

subroutine BICGSTAB(N, B, X, WORK, LDW, ITER, RESID, MATVEC, PSOLVE, INFO)
  use iso_c_binding
  use drd_gpu_c
  use drd_bicgstab_cb_m
  implicit none
  integer N, LDW, ITER, INFO
  double precision RESID
  double precision X(*), B(*), WORK(*)
  external MATVEC, PSOLVE
  integer(c_int32_t) :: it, inf
  real(c_double) :: res
  drd_user_matvec => MATVEC
  drd_user_psolve => PSOLVE
  it = ITER
  res = RESID
  inf = 0
  call drd_shim_check(drd_bicgstab_cb(c_null_ptr, int(N, c_int32_t), B, X, int(LDW, c_int32_t), it, res, &
       c_funloc(drd_cb_matvec), c_funloc(drd_cb_psolve), inf), 'BICGSTAB')
  ITER = it
  RESID = res
  INFO = inf
  drd_user_matvec => null()
  drd_user_psolve => null()
end subroutine BICGSTAB
