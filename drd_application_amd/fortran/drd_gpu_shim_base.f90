! drd_gpu_shim_base.f90 -- the device context of the TP drop-in (made from
! module `global`), its options, and the pressure projection PreSolve_init /
! PreSolve_2 (TP/PreSolve_init.f90, TP/PreSolve_2.f90) over the C-ABI of
! include/drd_gpu.h.  drd_gpu_shim.f90 adds the velocity and phi/mu solves;
! this file alone, with drd_gpu_bicgstab.f90, puts the pressure and the
! Krylov loops of the reference's own solve routines on the GPU.  Build it
! next to the caller's module `global` and drd_gpu_c.f90 with the caller's
! real kind.

! Creates the device context on first use (after finit/setup have set the
! module scalars) and refreshes its scalars on every later call.
subroutine drd_shim_ready()
  use global
  use drd_gpu_c
  implicit none
  type(drd_params_t) :: prm
  integer(c_int) :: rc
  call drd_params_default(prm, int(nr, c_int32_t), int(nz, c_int32_t))
  prm%phi_mode = phi_mode
  prm%lr_bd_type = ichar(lr_bd_type)
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
  prm%bdr_ph_used_adjust = bdr_ph_used_adjust
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


! Selects the fast path (no reference counterpart): velo_prec 3 = FP32
! fast-Poisson (DCT) velocity preconditioner, 4 = the same in FP64, 1 =
! identity, 2 = Jacobi, 0 = the reference's; ch_prec 1 = 2x2 block Jacobi
! for the phi/mu system, 0 = the reference's; krylov 1 = FGMRES(30) for the
! BDF2 velocity solve, 0 = BiCGSTAB.
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


! TP/PreSolve_init.f90:1-16.  FFTW plans have no meaning here: the context
! (and its DCT twiddles) is made at the first solve.
subroutine PreSolve_init(nr_, nz_, plan, plan_back)
  implicit none
  integer nr_, nz_
  integer*8 plan, plan_back
  plan = int(nr_, 8)
  plan_back = int(nz_, 8)
end subroutine PreSolve_init


! TP/PreSolve_2.f90:1-130
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
