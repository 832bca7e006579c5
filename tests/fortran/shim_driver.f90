! shim_driver.f90 -- test driver of the Fortran drop-in (drd_gpu_shim.f90).
!
! Linked against the caller-side module `global` (TP/global.f90, grid fixed
! at its compile-time nr x nz) exactly as a reference driver would be.  It
! reads module scalars and fields from a stream file written by the test,
! runs the reference's call sequence of nsteps time steps through the shim
! (TP/mcl6.f90:257-303 for the first, BDF1 step: ph_mu_solve,
! VeloSolve_combined, PreSolve_2; :395-502 for the BDF2 steps: ph_mu_solve,
! VeloSolve_2_combined, PreSolve_2, with the driver's own rotation of u_pre,
! ph_pre, rmu_pre, psi1 and p = psi2 + p), and writes the results to a
! second stream file.
!   shim_driver <in.bin> <out.bin> <nsteps>
program shim_driver
  use global
  implicit none
  real ph_after(-1:nr,-1:nz), rmu_after(-1:nr,-1:nz), psi1(-1:nr,-1:nz), psi2(-1:nr,-1:nz)
  real pf(-1:nr,-1:nz), ph_pre(-1:nr,-1:nz), rmu_pre(-1:nr,-1:nz)
  real u_after(-1:nr+1,-1:nz), v_after(-1:nr,0:nz), u_pre(-1:nr+1,-1:nz), v_pre(-1:nr,0:nz)
  real(8) scal(24)
  real t, gamma1
  integer*8 plan, plan_back
  integer :: uin, uout, i, j, s, nsteps
  character(len=1024) :: fin, fout, arg

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, arg)
  read(arg, *) nsteps
  call PreSolve_init(nr, nz, plan, plan_back)
  open(newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old')
  read(uin) scal
  read(uin) u, v, p, ph, rmu, rho, eta, bdr_ph, bdr_ph_used, bdr_ph_penalty, bdr_ph_grad, &
       ph_M_inner, angleB, angleT
  close(uin)
  dr = scal(1); dz = scal(2); dt = scal(3); re = scal(4); ca = scal(5); xld = scal(6)
  s_implicit = scal(7); I_thickness = scal(8); ggx = scal(9); ggz = scal(10)
  uwT = scal(11); uwB = scal(12); tol = scal(13); angle = scal(14); lambda_rho = scal(15)
  bdr_ph_used_adjust = scal(16); twall = scal(17); UU = scal(18); bdr_upp = scal(19)
  bdr_low = scal(20); xlsB = 0.0
  phi_mode = nint(scal(21))
  lr_bd_type = 'n'
  if (nint(scal(22)) == 1) lr_bd_type = 'd'
  dr2 = dr**2
  dz2 = dz**2
  do j = 0, nz
     do i = 0, nr
        r(i,j) = (i+0.5)*dr
        z(i,j) = (j+0.5)*dz
     end do
  end do
  pf = 0.0
  psi1 = 0.0
  psi2 = 0.0
  u_pre = u
  v_pre = v
  gamma1 = min(1.0, lambda_rho)
  t = 0.0
  do s = 1, nsteps
     if (s == 1) then
        call ph_mu_solve(ph_after, rmu_after, t)
        call VeloSolve_combined(psi1, u_after, v_after, rmu_after, ph_after, t+dt, pf, pf)
        u_pre = u
        v_pre = v
        u = u_after
        v = v_after
        call PreSolve_2(plan, plan_back, psi2, re*gamma1/dt)
        p = psi2 + p
        ph_pre = ph
        rmu_pre = rmu
        rmu = rmu_after
        ph = ph_after
     else
        call ph_mu_solve(ph_after, rmu_after, t+dt)
        call VeloSolve_2_combined(psi1, psi2, u_pre, u_after, v_pre, v_after, ph_pre, ph_after, &
             rmu_pre, rmu_after, t+dt, pf, pf)
        u_pre = u
        v_pre = v
        u = u_after
        v = v_after
        ph_pre = ph
        rmu_pre = rmu
        ph = ph_after
        rmu = rmu_after
        psi1 = psi2
        call PreSolve_2(plan, plan_back, psi2, 3*re*gamma1/2./dt)
        p = psi2 + p
     end if
     t = t + dt
  end do
  open(newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(uout) u, v, p, ph, rmu, psi2, ph_M_0, uL(0:nz-1), uR(0:nz-1)
  close(uout)
end program shim_driver
