! drd_gpu_shim_mf.f90 -- Fortran drop-in for the implicit-solve subroutines
! of the microfluidics program (microfluidics_complex_domain, "MF"), calling
! the C-ABI of include/drd_gpu.h through ISO_C_BINDING with the MF variant
! (DRD_VARIANT_MF: phi in [0, 1], the momentum right-hand side and operators
! of MF/VeloSolve_combined.f90 and MF/VeloSolve_2_combined.f90, PhSolve).
!
! Build it next to the caller's own module `global` (MF/global.f90) and the
! interface module drd_gpu_c.f90 with the caller's real kind (ifort -r8, or
! flang -fdefault-real-8 -fdefault-double-8), link libdrd_gpu.so, and leave
! the reference's VeloSolve_combined.f90, VeloSolve_2_combined.f90,
! PreSolve_init.f90, PreSolve_2.f90, dct_fftw.f90 and PhSolve.f90 off the
! link line.
!
! The subroutines keep the reference's names and dummy arguments
! (MF/VeloSolve_combined.f90:5-18, MF/VeloSolve_2_combined.f90:4-20,
! MF/PreSolve_init.f90:1-5, MF/PreSolve_2.f90:1-6, MF/PhSolve.f90:1-4).
! Hidden inputs are read from module `global` on every call; the hidden
! outputs uL, uR (the inflow profile, always on in MF) are written back.  A
! non-converged Krylov solve is reported and the run continues, as in the
! reference; a failure of the GPU library itself stops the program.

! Creates the device context on first use and refreshes its scalars later.
! MF's module has no phi_mode (its phase lies in [0, 1]), no
! bdr_ph_used_adjust and no separate penalty mask; its BDF2 momentum solve
! stops at 800 iterations (MF/VeloSolve_2_combined.f90:398), its PhSolve at
! 500 (MF/PhSolve.f90:72).
subroutine drd_shim_ready()
  use global
  use drd_gpu_c
  implicit none
  type(drd_params_t) :: prm
  integer(c_int) :: rc
  call drd_params_default(prm, int(nr, c_int32_t), int(nz, c_int32_t))
  prm%variant = 1
  prm%phi_mode = 1
  prm%lr_bd_type = ichar(lr_bd_type)
  prm%maxit_velo1 = 800
  prm%maxit_velo2 = 800
  prm%maxit_ph = 500
  prm%dx = dr
  prm%dz = dz
  prm%dt = dt
  prm%re = re
  prm%ca = ca
  prm%xld = xld
  prm%s_implicit = s_implicit
  prm%I_thickness = I_thickness
  prm%ggx = ggx
  prm%ggz = ggz
  prm%uwT = uwT
  prm%uwB = uwB
  prm%xlsB = xlsB
  prm%tol = tol
  prm%penalty_k = penalty_k
  prm%angle = angle
  prm%lambda_rho = lambda_rho
  prm%twall = twall
  prm%UU = UU
  prm%bdr_upp = bdr_upp
  prm%bdr_low = bdr_low
  prm%velo_prec = drd_shim_velo_prec
  prm%ch_prec = drd_shim_ch_prec
  prm%krylov = drd_shim_krylov
  if (.not. c_associated(drd_shim_ctx)) then
     rc = drd_init(prm, drd_shim_ctx)
     call drd_shim_check(rc, 'drd_init')
  else
     rc = drd_set_params(drd_shim_ctx, prm)
     call drd_shim_check(rc, 'drd_set_params')
  end if
end subroutine drd_shim_ready


! Selects the fast path (no reference counterpart), as in the TP shim.
subroutine drd_shim_options(velo_prec, ch_prec, krylov)
  use drd_gpu_c
  implicit none
  integer, intent(in) :: velo_prec, ch_prec, krylov
  drd_shim_velo_prec = int(velo_prec, c_int32_t)
  drd_shim_ch_prec = int(ch_prec, c_int32_t)
  drd_shim_krylov = int(krylov, c_int32_t)
end subroutine drd_shim_options


subroutine drd_shim_push(field, a)
  use drd_gpu_c
  implicit none
  integer(c_int32_t), intent(in) :: field
  real(c_double), intent(in) :: a(*)
  call drd_shim_check(drd_set_field(drd_shim_ctx, field, a), 'drd_set_field')
end subroutine drd_shim_push


! MF/PreSolve_init.f90:1-16 (FFTW plans have no meaning here)
subroutine PreSolve_init(nr_, nz_, plan, plan_back)
  implicit none
  integer nr_, nz_
  integer*8 plan, plan_back
  plan = int(nr_, 8)
  plan_back = int(nz_, 8)
end subroutine PreSolve_init


! MF/PreSolve_2.f90:1-130 (the same projection as TP's)
subroutine PreSolve_2(plan, plan_back, psi, coef)
  use global
  use drd_gpu_c
  implicit none
  integer*8 plan, plan_back
  real psi(-1:nr,-1:nz)
  real coef
  real(c_double) :: mean_div
  call drd_shim_ready()
  call drd_shim_push(DRD_F_U, u)
  call drd_shim_push(DRD_F_V, v)
  call drd_shim_check(drd_presolve_2(drd_shim_ctx, psi, real(coef, c_double), mean_div), 'PreSolve_2')
  print *, 'output SUM SUM', mean_div
end subroutine PreSolve_2


! hidden inflow outputs uL, uR of MF/VeloSolve_2_combined.f90:261-295
subroutine drd_shim_inflow(t)
  use global
  implicit none
  real t, cf, tmp_hig
  integer k
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


! the fields the MF momentum solves read: the smooth bdr_ph is both the
! surface-tension weight and the penalty mask
subroutine drd_shim_push_velo()
  use global
  use drd_gpu_c
  implicit none
  call drd_shim_push(DRD_F_U, u)
  call drd_shim_push(DRD_F_V, v)
  call drd_shim_push(DRD_F_P, p)
  call drd_shim_push(DRD_F_RHO, rho)
  call drd_shim_push(DRD_F_ETA, eta)
  call drd_shim_push(DRD_F_BDR_PH, bdr_ph)
end subroutine drd_shim_push_velo


! MF/VeloSolve_combined.f90:5-532
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


! MF/VeloSolve_2_combined.f90:4-510
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


! MF/PhSolve.f90:1-126: the mu-only CH solve (DCT-I preconditioner) and the
! phi recovery; ph_M_0 and ph_M_inner are the caller's (static in MF).
subroutine PhSolve(rmu_after, ph_after, t)
  use global
  use drd_gpu_c
  implicit none
  real rmu_after(-1:nr,-1:nz), ph_after(-1:nr,-1:nz), t
  type(drd_solve_info_t) :: info
  call drd_shim_ready()
  call drd_shim_push(DRD_F_U, u)
  call drd_shim_push(DRD_F_V, v)
  call drd_shim_push(DRD_F_PH, ph)
  call drd_shim_push(DRD_F_RMU, rmu)
  call drd_shim_push(DRD_F_BDR_PH, bdr_ph)
  call drd_shim_push(DRD_F_BDR_PH_GRAD, bdr_ph_grad)
  call drd_shim_push(DRD_F_PH_M_INNER, ph_M_inner)
  call drd_shim_push(DRD_F_PH_M_0, ph_M_0)
  call drd_shim_check(drd_phsolve(drd_shim_ctx, rmu_after, ph_after, real(t, c_double), info), 'PhSolve')
  write(*,*) '\mu BICGSTAB', info%iter, info%resid
  if (info%info /= 0) print *, 'error', info%info, tol
end subroutine PhSolve
