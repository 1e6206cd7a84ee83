! Writes one wavefunction snapshot with the Fortran runtime's own list-directed
! and unformatted sequential output, record for record as printoutc.f90:69-114
! does (two comment lines, then grid, rimp, vimp, ninvar, invar, psi).  The
! values are seeded and include runs of equal values (repeat counts "k*").
program snapwrite
  implicit none
  integer, parameter :: nx = 8, ny = 6, nz = 5
  complex(8) :: psi(nx, ny, nz), invar(3)
  real(8) :: rimp(3), vimp(3), x
  integer :: ix, iy, iz
  character(len=256) :: fmtname, binname
  call get_command_argument(1, fmtname)
  call get_command_argument(2, binname)
  rimp = (/ 0.125d0, -1.5d0, 2.0d-3 /)
  vimp = (/ 0.d0, 0.d0, 0.d0 /)
  invar = (/ (1.d0, 0.d0), (0.d0, -0.5d0), (3.25d0, 1.d-7) /)
  do iz = 1, nz
    do iy = 1, ny
      do ix = 1, nx
        x = sin(0.37d0 * ix + 1.3d0 * iy - 0.71d0 * iz) * exp(-0.05d0 * (ix + iy + iz))
        psi(ix, iy, iz) = cmplx(x, 0.1d0 * x * x, kind=8)
        if (iz == 3) psi(ix, iy, iz) = (0.d0, 0.d0)
      end do
    end do
  end do
  open(10, file=trim(fmtname), form='FORMATTED')
  write(10, '("# snapshot written by the Fortran runtime")')
  write(10, '("#  xcm,ycm,zcm:",1p,3E16.8)') 0.d0, 0.d0, 0.d0
  write(10, *) 4.d0, 3.d0, 2.5d0, 1.d0, 1.d0, 1.d0, nx, ny, nz
  write(10, *) rimp
  write(10, *) vimp
  write(10, *) 3
  write(10, *) invar
  write(10, *) psi
  close(10)
  open(11, file=trim(binname), form='UNFORMATTED')
  write(11) 4.d0, 3.d0, 2.5d0, 1.d0, 1.d0, 1.d0, nx, ny, nz
  write(11) rimp
  write(11) vimp
  write(11) 3
  write(11) invar
  write(11) psi
  close(11)
end program snapwrite
