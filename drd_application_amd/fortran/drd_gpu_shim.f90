! drd_gpu_shim.f90 -- Fortran drop-in for the reference's implicit-solve
! subroutines, calling the C-ABI of include/drd_gpu.h through ISO_C_BINDING.
!
! Build it and drd_gpu_shim_base.f90 (the context, PreSolve_init and
! PreSolve_2) next to the caller's own module `global` (TP/global.f90) and
! the interface module drd_gpu_c.f90 with the caller's real kind (ifort -r8,
! or flang -fdefault-real-8 -fdefault-double-8), link libdrd_gpu.so, and
! leave the reference's VeloSolve_combined.f90, VeloSolve_2_combined.f90,
! PreSolve_init.f90, PreSolve_2.f90, dct_fftw.f90 and ph_mu_solve.f90 off
! the link line.  The MF program takes drd_gpu_shim_mf.f90 instead.
!
! The subroutines keep the reference's names and dummy arguments.  Hidden
! inputs are read from module `global` and copied to the device context on
! every call; hidden outputs (uL, uR, coef_bot, coef_top, ph_M_0) are
! written back.  Iteration counts are printed as the reference prints them,
! and a non-converged Krylov solve is reported and the run continues, as in
! the reference; a failure of the GPU library itself stops the program.

! hidden inflow outputs uL, uR of TP/VeloSolve_combined.f90:266-306
subroutine drd_shim_inflow(t)
  use global
  implicit none
  real t, cf, tmp_hig
  integer k
  if (lr_bd_type /= 'd') return
  cf = 1.0 - exp(-t/twall)
  tmp_hig = bdr_upp - bdr_low
  do k = 0, nz-1
     if (z(0,k) > bdr_low .and. z(0,k) < bdr_upp) then
        uL(k) = UU*( -(z(0,k) - bdr_low - tmp_hig/2)**2/tmp_hig/tmp_hig*4 + 1 )*cf
     else
        uL(k) = 0.0
     end if
     uR(k) = uL(k)
  end do
end subroutine drd_shim_inflow


subroutine drd_shim_push_velo()
  use global
  use drd_gpu_c
  implicit none
  call drd_shim_push(DRD_F_U, u)
  call drd_shim_push(DRD_F_V, v)
  call drd_shim_push(DRD_F_P, p)
  call drd_shim_push(DRD_F_RHO, rho)
  call drd_shim_push(DRD_F_ETA, eta)
  call drd_shim_push(DRD_F_BDR_PH_USED, bdr_ph_used)
  call drd_shim_push(DRD_F_BDR_PH_PENALTY, bdr_ph_penalty)
end subroutine drd_shim_push_velo


! TP/VeloSolve_combined.f90:5-532
subroutine VeloSolve_combined(psi, u_after, v_after, rmu_after, ph_after, t, &
     particle_force1, particle_force2)
  use global
  use drd_gpu_c
  implicit none
  real psi(-1:nr,-1:nz)
  real u_after(-1:nr+1,-1:nz), v_after(-1:nr,0:nz)
  real rmu_after(-1:nr,-1:nz), ph_after(-1:nr,-1:nz), t
  real particle_force1(-1:nr,-1:nz), particle_force2(-1:nr,-1:nz)
  type(drd_solve_info_t) :: info
  call drd_shim_ready()
  call drd_shim_push_velo()
  call drd_shim_check(drd_velosolve_combined(drd_shim_ctx, psi, u_after, v_after, rmu_after, &
       ph_after, real(t, c_double), particle_force1, particle_force2, info), 'VeloSolve_combined')
  call drd_shim_inflow(t)
  write(*,*) 'uv 1 BICGSTAB no slip', info%iter
  if (info%info /= 0) print *, 'uv 1 BICGSTAB no slip: error', info%iter, info%info
end subroutine VeloSolve_combined


! TP/VeloSolve_2_combined.f90:4-531
subroutine VeloSolve_2_combined(psi1, psi2, u_pre, u_after, v_pre, v_after, &
     ph_pre, ph_after, rmu_pre, rmu_after, t, particle_force1, particle_force2)
  use global
  use drd_gpu_c
  implicit none
  real psi1(-1:nr,-1:nz), psi2(-1:nr,-1:nz)
  real u_pre(-1:nr+1,-1:nz), u_after(-1:nr+1,-1:nz)
  real v_pre(-1:nr,0:nz), v_after(-1:nr,0:nz)
  real rmu_pre(-1:nr,-1:nz), rmu_after(-1:nr,-1:nz)
  real ph_pre(-1:nr,-1:nz), ph_after(-1:nr,-1:nz), t
  real particle_force1(-1:nr,-1:nz), particle_force2(-1:nr,-1:nz)
  type(drd_solve_info_t) :: info
  call drd_shim_ready()
  call drd_shim_push_velo()
  call drd_shim_check(drd_velosolve_2_combined(drd_shim_ctx, psi1, psi2, u_pre, u_after, v_pre, &
       v_after, ph_pre, ph_after, rmu_pre, rmu_after, real(t, c_double), particle_force1, &
       particle_force2, info), 'VeloSolve_2_combined')
  call drd_shim_inflow(t)
  write(*,*) 'uv 2 BICGSTAB no slip', info%iter
  if (info%info /= 0) print *, 'uv 2 BICGSTAB no slip: error', info%iter, info%info
end subroutine VeloSolve_2_combined


! TP/ph_mu_solve.f90:3-197
subroutine ph_mu_solve(ph_after, rmu_after, t)
  use global
  use drd_gpu_c
  implicit none
  real ph_after(-1:nr,-1:nz), rmu_after(-1:nr,-1:nz), t
  type(drd_solve_info_t) :: info
  call drd_shim_ready()
  call drd_shim_push(DRD_F_U, u)
  call drd_shim_push(DRD_F_V, v)
  call drd_shim_push(DRD_F_PH, ph)
  call drd_shim_push(DRD_F_RMU, rmu)
  call drd_shim_push(DRD_F_BDR_PH, bdr_ph)
  call drd_shim_push(DRD_F_BDR_PH_USED, bdr_ph_used)
  call drd_shim_push(DRD_F_BDR_PH_GRAD, bdr_ph_grad)
  call drd_shim_push(DRD_F_PH_M_INNER, ph_M_inner)
  call drd_shim_push(DRD_F_ANGLEB, angleB)
  call drd_shim_push(DRD_F_ANGLET, angleT)
  call drd_shim_check(drd_ph_mu_solve(drd_shim_ctx, ph_after, rmu_after, real(t, c_double), info), &
       'ph_mu_solve')
  ! hidden outputs: ph_M_0 (:31-35) and the wall coefficients (:123, :144)
  call drd_shim_check(drd_get_field(drd_shim_ctx, DRD_F_PH_M_0, ph_M_0), 'drd_get_field')
  coef_bot = (I_thickness/dz) / (I_thickness/dz)
  coef_top = (I_thickness/dz) / (I_thickness/dz)
  write(*,*) 'ph_mu_solve: \phi 1 BICGSTAB', info%iter, info%resid, tol
  if (info%info /= 0) print *, 'error', info%info, tol
end subroutine ph_mu_solve

