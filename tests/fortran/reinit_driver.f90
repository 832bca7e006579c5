! reinit_driver.f90 -- test driver of the reference's own CH re-initialisation
! of a solid domain (TP/reinint_by_PhSolve.f90, with PhSolve_ori.f90 and
! setup.f90 compiled from their sources) over the Krylov drop-in BICGSTAB of
! drd_gpu_bicgstab.f90.
!
! Reads module scalars and fields from a stream file written by the test
! (scal(8) = dr, dz, xld, I_thickness, s_implicit, tol, alpha_s, re; then
! bdr_ph, u, v, angleB, angleT), calls reinint_by_PhSolve, and writes
! bdr_ph, bdr_ph_penalty, ph, rmu.
!   reinit_driver <in.bin> <out.bin>
program reinit_driver
  use global
  implicit none
  real(8) scal(8)
  integer :: uin, uout, i, j
  character(len=1024) :: fin, fout

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old')
  read(uin) scal
  read(uin) bdr_ph, u, v, angleB, angleT
  close(uin)
  dr = scal(1); dz = scal(2); xld = scal(3); I_thickness = scal(4); s_implicit = scal(5)
  tol = scal(6); alpha_s = scal(7); re = scal(8)
  dr2 = dr**2
  dz2 = dz**2
  lambda_rho = 1.0; lambda_eta = 1.0; slipratio = 1.0; xlsB = 0.0; xlsT = 0.0
  lr_bd_type = 'n'; top_type_ph = 'n'; bot_type_ph = 'n'
  do j = 0, nz
     do i = 0, nr
        r(i,j) = (i+0.5)*dr
        z(i,j) = (j+0.5)*dz
     end do
  end do
  call reinint_by_PhSolve()
  open(newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(uout) bdr_ph, bdr_ph_penalty, ph, rmu
  close(uout)
end program reinit_driver
