!-----------------------------------------------------------------------
! interp_gpu -- the reference's interpolation tool (interpolate/main.F90)
! on the MI355X library: reads restart.nc, doubles the horizontal
! resolution of R, T, U, V, W, Bx, By, Bz by spectral zero-padding
! (expand, interpolate/main.F90:299-306) and writes output.nc in FP64 with
! dimensions (2nx, 2ny, nz) and the restart's time, iteration, time step
! and parameters (interpolate/main.F90:371-399).
!-----------------------------------------------------------------------
program interp_gpu
  use iso_c_binding
  use mhd_capi
  implicit none

  type(mhd_dump_t) :: dmp
  integer(c_int)   :: ier
  integer          :: f, nx, ny, nz
  real(c_double), allocatable, target :: coarse(:,:,:,:), fine(:,:,:,:)
  type(c_ptr)      :: cp(8), fp(8)
  real(c_double)   :: mn, mx
  character(len=2), parameter :: names(8) = ['R ', 'T ', 'U ', 'V ', 'W ', 'Bx', 'By', 'Bz']

  ier = mhd_read_dump('restart.nc'//c_null_char, dmp, [(c_null_ptr, f = 1, 8)])
  if (ier /= 0) call die(ier, 'reading restart.nc')
  nx = dmp%nx; ny = dmp%ny; nz = dmp%nz
  print *, ' nx ', nx, ' ny ', ny, ' nz ', nz, ' ->', 2*nx, 2*ny, nz
  allocate(coarse(nx, ny, nz, 8), fine(2*nx, 2*ny, nz, 8))
  do f = 1, 8
    cp(f) = c_loc(coarse(1, 1, 1, f))
    fp(f) = c_loc(fine(1, 1, 1, f))
  end do
  ier = mhd_read_dump('restart.nc'//c_null_char, dmp, cp)
  if (ier /= 0) call die(ier, 'reading restart.nc')

  do f = 1, 8
    ier = mhd_expand(nx, ny, nz, coarse(:, :, :, f), fine(:, :, :, f))
    if (ier /= 0) call die(ier, 'expand')
    mn = minval(fine(:, :, :, f)); mx = maxval(fine(:, :, :, f))
    print *, ' ', names(f), 'min= ', mn, ' max= ', mx
  end do

  dmp%nx = 2*nx
  dmp%ny = 2*ny
  ier = mhd_write_dump('output.nc'//c_null_char, dmp, fp, 8)
  if (ier /= 0) call die(14_c_int, 'writing output.nc')
  print *, ' wrote output.nc'

contains

  subroutine die(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    print *, 'ERROR: ', what, ': ', trim(mhd_message(code)), ' (code ', code, ')'
    error stop 1
  end subroutine

end program interp_gpu
