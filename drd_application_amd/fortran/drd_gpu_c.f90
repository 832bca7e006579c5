! drd_gpu_c.f90 -- ISO_C_BINDING interfaces of the C-ABI (include/drd_gpu.h)
! that the Fortran drop-in shims call: drd_gpu_shim.f90 (the TP program,
! two_phase_flow_complex_domain), drd_gpu_shim_mf.f90 (the MF program,
! microfluidics_complex_domain) and drd_gpu_bicgstab.f90 (BICGSTAB itself).
! It uses no module `global`, so one build serves all of them.

module drd_gpu_c
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: DRD_F_U = 0, DRD_F_V = 1, DRD_F_P = 2, DRD_F_PH = 3, &
       DRD_F_RMU = 4, DRD_F_RHO = 5, DRD_F_ETA = 6, DRD_F_BDR_PH = 7, DRD_F_BDR_PH_USED = 8, &
       DRD_F_BDR_PH_PENALTY = 9, DRD_F_BDR_PH_GRAD = 10, DRD_F_PH_M_INNER = 11, &
       DRD_F_PH_M_0 = 12, DRD_F_ANGLEB = 13, DRD_F_ANGLET = 14

  ! drd_params (include/drd_gpu.h): 12 int32 then doubles, no padding
  type, bind(C) :: drd_params_t
     integer(c_int32_t) :: nx, ny, variant, phi_mode, lr_bd_type, velo_prec
     integer(c_int32_t) :: maxit_velo1, maxit_velo2, maxit_phmu, maxit_ph, ch_prec, krylov
     real(c_double) :: dx, dz, dt, re, ca, xld, s_implicit, I_thickness
     real(c_double) :: ggx, ggz, uwT, uwB, xlsB, tol, penalty_k, angle, lambda_rho
     real(c_double) :: bdr_ph_used_adjust, twall, UU, bdr_upp, bdr_low
  end type drd_params_t

  type, bind(C) :: drd_solve_info_t
     integer(c_int32_t) :: iter, info
     real(c_double) :: resid, true_resid
  end type drd_solve_info_t

  interface
     subroutine drd_params_default(p, nx, ny) bind(C, name='drd_params_default')
       import :: drd_params_t, c_int32_t
       type(drd_params_t), intent(out) :: p
       integer(c_int32_t), value :: nx, ny
     end subroutine drd_params_default

     integer(c_int) function drd_init(p, ctx) bind(C, name='drd_init')
       import :: drd_params_t, c_ptr, c_int
       type(drd_params_t), intent(in) :: p
       type(c_ptr), intent(out) :: ctx
     end function drd_init

     integer(c_int) function drd_set_params(ctx, p) bind(C, name='drd_set_params')
       import :: drd_params_t, c_ptr, c_int
       type(c_ptr), value :: ctx
       type(drd_params_t), intent(in) :: p
     end function drd_set_params

     integer(c_int) function drd_set_field(ctx, field, host) bind(C, name='drd_set_field')
       import :: c_ptr, c_int, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: field
       real(c_double), intent(in) :: host(*)
     end function drd_set_field

     integer(c_int) function drd_get_field(ctx, field, host) bind(C, name='drd_get_field')
       import :: c_ptr, c_int, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: field
       real(c_double), intent(out) :: host(*)
     end function drd_get_field

     integer(c_int) function drd_velosolve_combined(ctx, psi, u_after, v_after, rmu_after, &
          ph_after, t, pf1, pf2, info) bind(C, name='drd_velosolve_combined')
       import :: c_ptr, c_int, c_double, drd_solve_info_t
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: psi(*), rmu_after(*), ph_after(*), pf1(*), pf2(*)
       real(c_double), intent(out) :: u_after(*), v_after(*)
       real(c_double), value :: t
       type(drd_solve_info_t), intent(out) :: info
     end function drd_velosolve_combined

     integer(c_int) function drd_velosolve_2_combined(ctx, psi1, psi2, u_pre, u_after, v_pre, &
          v_after, ph_pre, ph_after, rmu_pre, rmu_after, t, pf1, pf2, info) &
          bind(C, name='drd_velosolve_2_combined')
       import :: c_ptr, c_int, c_double, drd_solve_info_t
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: psi1(*), psi2(*), u_pre(*), v_pre(*), ph_pre(*), ph_after(*)
       real(c_double), intent(in) :: rmu_pre(*), rmu_after(*), pf1(*), pf2(*)
       real(c_double), intent(out) :: u_after(*), v_after(*)
       real(c_double), value :: t
       type(drd_solve_info_t), intent(out) :: info
     end function drd_velosolve_2_combined

     integer(c_int) function drd_presolve_2(ctx, psi, coef, mean_div) bind(C, name='drd_presolve_2')
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: psi(*)
       real(c_double), value :: coef
       real(c_double), intent(out) :: mean_div
     end function drd_presolve_2

     integer(c_int) function drd_ph_mu_solve(ctx, ph_after, rmu_after, t, info) &
          bind(C, name='drd_ph_mu_solve')
       import :: c_ptr, c_int, c_double, drd_solve_info_t
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: ph_after(*), rmu_after(*)
       real(c_double), value :: t
       type(drd_solve_info_t), intent(out) :: info
     end function drd_ph_mu_solve

     integer(c_int) function drd_phsolve(ctx, rmu_after, ph_after, t, info) bind(C, name='drd_phsolve')
       import :: c_ptr, c_int, c_double, drd_solve_info_t
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: rmu_after(*), ph_after(*)
       real(c_double), value :: t
       type(drd_solve_info_t), intent(out) :: info
     end function drd_phsolve

     integer(c_int) function drd_bicgstab_cb(ctx, n, b, x, ldw, iter, resid, matvec, psolve, info) &
          bind(C, name='drd_bicgstab_cb')
       import :: c_ptr, c_int, c_int32_t, c_double, c_funptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n, ldw
       real(c_double), intent(in) :: b(*)
       real(c_double), intent(inout) :: x(*)
       integer(c_int32_t), intent(inout) :: iter, info
       real(c_double), intent(inout) :: resid
       type(c_funptr), value :: matvec, psolve
     end function drd_bicgstab_cb

     type(c_ptr) function drd_last_error() bind(C, name='drd_last_error')
       import :: c_ptr
     end function drd_last_error
  end interface

  type(c_ptr), save :: drd_shim_ctx = c_null_ptr
  ! solver options (drd_params velo_prec, ch_prec, krylov); 0, 0, 0 is the
  ! reference's choice.  Set with drd_shim_options before the first solve.
  integer(c_int32_t), save :: drd_shim_velo_prec = 0, drd_shim_ch_prec = 0, drd_shim_krylov = 0

contains

  ! stops the program with the library's message when a call failed
  subroutine drd_shim_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=256) :: text
    integer :: i
    if (rc == 0) return
    call c_f_pointer(drd_last_error(), msg, [256])
    text = ' '
    do i = 1, 256
       if (msg(i) == c_null_char) exit
       text(i:i) = msg(i)
    end do
    write(*, '(a,a,a,i0,a,a)') 'drd_gpu: ', where, ' failed (', rc, '): ', trim(text)
    error stop 1
  end subroutine drd_shim_check

end module drd_gpu_c
