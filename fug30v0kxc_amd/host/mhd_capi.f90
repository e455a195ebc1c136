!-----------------------------------------------------------------------
! mhd_capi -- iso_c_binding interface to the C ABI of include/mhd.h
! (libmhd_hip.so, the MI355X time-step library).  Derived types mirror the
! C structs field by field; every function returns the C ABI's status code
! (0 = ok, otherwise the reference's abort / iret codes).
!-----------------------------------------------------------------------
module mhd_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: MHD_PREC_FP64 = 0, MHD_PREC_MIXED = 1
  integer(c_int), parameter :: MHD_OPT_RADBC = 1, MHD_OPT_CONSTFLUX = 2, MHD_OPT_PERFECT = 4, &
                               MHD_OPT_KINEMATIC = 8, MHD_OPT_DYNAMO = 16
  integer(c_int), parameter :: MHD_F_U = 0, MHD_F_V = 1, MHD_F_W = 2, MHD_F_T = 3, &
                               MHD_F_R = 4, MHD_F_BX = 5, MHD_F_BY = 6, MHD_F_BZ = 7

  type, bind(C) :: mhd_params_t
    real(c_double) :: gamma, sigma, theta, ck, cm, f, tayl, psi, zeta, xmax, ymax, sf
    real(c_double) :: t_amp, b_amp
    real(c_double) :: eta_c = 0.0d0, eta_z0 = 0.5d0, eta_w = 0.05d0
    integer(c_int) :: opts = 0
  end type

  type, bind(C) :: mhd_scalars_t
    real(c_double) :: dt, dt1, dt2, ttime, bzbar
    integer(c_int) :: nit, iter, first, second
  end type

  type, bind(C) :: mhd_diag_t
    real(c_double) :: ubar, vbar, kinenergy, magenergy, nusselt, divbmax, divbmin, dt
    integer(c_int) :: nit
  end type

  type, bind(C) :: mhd_dump_t
    integer(c_int) :: nx, ny, nz, version, iteration
    real(c_double) :: time, timestep
    real(c_double) :: gamma, sigma, theta, ck, cm, f, tayl, psi, zeta, xmax, ymax
    character(kind=c_char) :: history(256)
  end type

  interface
    subroutine mhd_ra_q_to_ck_f(gamma, sigma, theta, cm, zeta, ra, q, ck, f) bind(C)
      import :: c_double
      real(c_double), value :: gamma, sigma, theta, cm, zeta, ra, q
      real(c_double) :: ck, f
    end subroutine

    integer(c_int) function mhd_create(p, nx, ny, nz, prec, ctx) bind(C)
      import :: c_int, c_ptr, mhd_params_t
      type(mhd_params_t) :: p
      integer(c_int), value :: nx, ny, nz, prec
      type(c_ptr) :: ctx
    end function

    subroutine mhd_destroy(ctx) bind(C)
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine

    integer(c_int) function mhd_static(ctx) bind(C)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function mhd_set_state(ctx, r, t, u, v, w, bx, by, bz) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: r(*), t(*), u(*), v(*), w(*), bx(*), by(*), bz(*)
    end function

    integer(c_int) function mhd_get_state(ctx, r, t, u, v, w, bx, by, bz) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: r(*), t(*), u(*), v(*), w(*), bx(*), by(*), bz(*)
    end function

    integer(c_int) function mhd_get_scalars(ctx, s) bind(C)
      import :: c_int, c_ptr, mhd_scalars_t
      type(c_ptr), value :: ctx
      type(mhd_scalars_t) :: s
    end function

    integer(c_int) function mhd_set_scalars(ctx, s) bind(C)
      import :: c_int, c_ptr, mhd_scalars_t
      type(c_ptr), value :: ctx
      type(mhd_scalars_t) :: s
    end function

    integer(c_int) function mhd_step(ctx, nsteps, dt_fixed, ndu06, last) bind(C)
      import :: c_int, c_ptr, c_double, mhd_diag_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps, ndu06
      real(c_double), value :: dt_fixed
      type(mhd_diag_t) :: last
    end function

    integer(c_int) function mhd_sminmax(ctx, id, mm) bind(C)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: id
      real(c_double) :: mm(2)
    end function

    integer(c_int) function mhd_write_dump(path, meta, fields, iprec) bind(C)
      import :: c_int, c_ptr, c_char, mhd_dump_t
      character(kind=c_char) :: path(*)
      type(mhd_dump_t) :: meta
      type(c_ptr) :: fields(8)
      integer(c_int), value :: iprec
    end function

    integer(c_int) function mhd_read_dump(path, meta, fields) bind(C)
      import :: c_int, c_ptr, c_char, mhd_dump_t
      character(kind=c_char) :: path(*)
      type(mhd_dump_t) :: meta
      type(c_ptr) :: fields(8)
    end function

    integer(c_int) function mhd_expand(nx, ny, nz, fin, fout) bind(C)
      import :: c_int, c_double
      integer(c_int), value :: nx, ny, nz
      real(c_double) :: fin(*), fout(*)
    end function

    integer(c_int) function mhd_box_filter(nx, ny, nz, f, fin, fout) bind(C)
      import :: c_int, c_double
      integer(c_int), value :: nx, ny, nz, f
      real(c_double) :: fin(*), fout(*)
    end function

    type(c_ptr) function mhd_strerror(code) bind(C)
      import :: c_int, c_ptr
      integer(c_int), value :: code
    end function
  end interface

contains

  ! C string of mhd_strerror(code) as a Fortran string
  function mhd_message(code) result(msg)
    integer(c_int), intent(in) :: code
    character(len=64) :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: i
    call c_f_pointer(mhd_strerror(code), s, [64])
    msg = ' '
    do i = 1, 64
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function

end module mhd_capi
