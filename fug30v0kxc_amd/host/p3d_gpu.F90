!-----------------------------------------------------------------------
! p3d_gpu -- host driver of the MI355X time-step library.
!
! Keeps the reference program's interface for a user: the 25-line INPUT
! file (src/initial.F90:142-171, flag 1 = CK/F, flag 2 = Ra/Q), the time
! loop cadence of src/main.F90:340-747 (diagnostics every NDU06 steps,
! T16_XXX.nc dumps every NDU16 steps, FP64 dumps at the end and every
! NDURESTART dumps, FP32 otherwise), restarts from restart.nc (NFRAME 1:
! parameters from the file, NFRAME 2: from INPUT) and the abort codes.
! Every per-step routine runs on the GPU behind the C ABI (mhd_capi).
!
! Optional extension file INPUT.ext, lines "key value":
!   dt_fixed  <real>   fixed time step (default 0: the reference's calcdt)
!   t_amp     <real>   T noise amplitude of the static start (default 0.1)
!   b_amp     <real>   Bz noise amplitude (default 0)
!   precision fp64|mixed
!   eta_c, eta_z0, eta_w  <real>  eta(z) step: contrast, centre, width
!             (extension; contrast <= 1 keeps the reference's constant eta)
!   radbc | constflux | perfect | kinematic | dynamo   1 to switch on the
!             reference's compile-time option (src/main.F90:109-116)
!-----------------------------------------------------------------------
program p3d_gpu
  use iso_c_binding
  use mhd_capi
  implicit none

  type(mhd_params_t) :: prm
  type(mhd_diag_t)   :: diag
  type(mhd_scalars_t) :: sc
  type(mhd_dump_t)   :: dmp
  type(c_ptr)        :: ctx
  integer(c_int)     :: ier, prec, iprec
  integer            :: flag, nframe, ntotal, ndu06, ndu16, ndurestart
  integer            :: nx, ny, nz, nit, idump
  real(c_double)     :: dt_fixed, rayl, chand, mm(2)
  real(c_double), allocatable, target :: r(:,:,:), t(:,:,:), u(:,:,:), v(:,:,:), &
                                         w(:,:,:), bx(:,:,:), by(:,:,:), bz(:,:,:)
  type(c_ptr)        :: fp(8)
  integer(kind=8)    :: clk0, clk1, rate
  character(len=16)  :: fname
  character(len=80)  :: hist

  ctx = c_null_ptr
  prm%t_amp = 0.1d0          ! src/initial.F90:31
  prm%b_amp = 0.0d0
  dt_fixed = 0.0d0
  prec = MHD_PREC_FP64

  call read_input('INPUT', flag, nframe, ntotal, ndu06, ndu16, ndurestart, nx, ny, nz)
  call read_extensions('INPUT.ext')

  if (nframe > 0) then
    ! dimensions, time, iteration and (NFRAME 1) parameters from restart.nc
    ier = mhd_read_dump('restart.nc'//c_null_char, dmp, [(c_null_ptr, idump = 1, 8)])
    if (ier /= 0) call die(ier, 'reading restart.nc')
    if (nframe == 1) then
      prm%gamma = dmp%gamma; prm%sigma = dmp%sigma; prm%theta = dmp%theta
      prm%ck = dmp%ck; prm%cm = dmp%cm; prm%f = dmp%f; prm%tayl = dmp%tayl
      prm%psi = dmp%psi; prm%zeta = dmp%zeta; prm%xmax = dmp%xmax; prm%ymax = dmp%ymax
    end if
    nx = dmp%nx; ny = dmp%ny; nz = dmp%nz
  end if

  print *, '---------------------------------------------------'
  print *, 'Simulations of compressible 3-D magnetoconvection  '
  print *, '  MI355X time-step library, ', trim(merge('FP64 ', 'mixed', prec == MHD_PREC_FP64))
  print *, '---------------------------------------------------'
  print *, ' nx = ', nx, ' ny = ', ny, ' nz = ', nz
  print *, ' GAMMA = ', prm%gamma, ' THETA = ', prm%theta, ' CK = ', prm%ck
  print *, ' F     = ', prm%f, ' ZETA = ', prm%zeta, ' SF = ', prm%sf

  ier = mhd_create(prm, nx, ny, nz, prec, ctx)
  if (ier /= 0) call die(ier, 'mhd_create')

  allocate(r(nx,ny,nz), t(nx,ny,nz), u(nx,ny,nz), v(nx,ny,nz), w(nx,ny,nz), &
           bx(nx,ny,nz), by(nx,ny,nz), bz(nx,ny,nz))
  fp = [c_loc(r), c_loc(t), c_loc(u), c_loc(v), c_loc(w), c_loc(bx), c_loc(by), c_loc(bz)]

  if (nframe == 0) then
    ier = mhd_static(ctx)
    if (ier /= 0) call die(ier, 'static')
  else
    ier = mhd_read_dump('restart.nc'//c_null_char, dmp, fp)
    if (ier /= 0) call die(ier, 'reading restart.nc')
    ier = mhd_set_state(ctx, r, t, u, v, w, bx, by, bz)
    if (ier /= 0) call die(ier, 'set_state')
    ! restarts begin with an Euler step (src/main.F90:110, 658-661)
    ier = mhd_get_scalars(ctx, sc)
    sc%iter = dmp%iteration
    sc%ttime = dmp%time
    sc%dt1 = dmp%timestep
    sc%first = 1
    sc%second = 0
    ier = mhd_set_scalars(ctx, sc)
    print *, 'RESTART ttime ', dmp%time, ' iter ', dmp%iteration
  end if

  call system_clock(clk0, rate)
  do nit = 1, ntotal
    ier = mhd_step(ctx, 1, dt_fixed, ndu06, diag)
    if (ier /= 0) call die(ier, 'time step')

    if (mod(nit, ndu06) == 0) then
      ier = mhd_get_scalars(ctx, sc)
      print *
      print *, ' NIT= ', nit, ' TTIME ', sc%ttime
      print *, ' ubar= ', diag%ubar, ' vbar= ', diag%vbar
      print *, ' divbmin= ', diag%divbmin, ' divbmax= ', diag%divbmax
      ier = mhd_sminmax(ctx, MHD_F_R, mm); print *, ' rmin= ', mm(1), ' rmax= ', mm(2)
      ier = mhd_sminmax(ctx, MHD_F_T, mm); print *, ' tmin= ', mm(1), ' tmax= ', mm(2)
      ier = mhd_sminmax(ctx, MHD_F_W, mm); print *, ' wmin= ', mm(1), ' wmax= ', mm(2)
      ier = mhd_sminmax(ctx, MHD_F_BZ, mm); print *, ' bzmin= ', mm(1), ' bzmax= ', mm(2)
      print *, ' magenergy= ', diag%magenergy
      print *, ' kinenergy= ', diag%kinenergy
      print *, ' Nusselt= ', diag%nusselt
      print *, ' dt= ', diag%dt
    end if

    if (mod(nit, ndu16) == 0) then
      ! src/main.F90:695-740: FP32 dumps, FP64 at the end and every NDURESTART
      iprec = 4
      if (nit >= ntotal) iprec = 8
      if (mod(nit, ndurestart*ndu16) == 0) iprec = 8
      ier = mhd_get_scalars(ctx, sc)
      idump = mod(sc%iter/ndu16, 1000)
      write(fname, '(a,i3.3,a)') 'T16_', idump, '.nc'
      ier = mhd_get_state(ctx, r, t, u, v, w, bx, by, bz)
      if (ier /= 0) call die(ier, 'get_state')
      dmp%nx = nx; dmp%ny = ny; dmp%nz = nz
      dmp%version = 3
      dmp%iteration = sc%iter
      dmp%time = sc%ttime
      dmp%timestep = sc%dt1
      dmp%gamma = prm%gamma; dmp%sigma = prm%sigma; dmp%theta = prm%theta
      dmp%ck = prm%ck; dmp%cm = prm%cm; dmp%f = prm%f; dmp%tayl = prm%tayl
      dmp%psi = prm%psi; dmp%zeta = prm%zeta; dmp%xmax = prm%xmax; dmp%ymax = prm%ymax
      hist = 'Created by p3d_gpu (MI355X time-step library)'
      call set_history(dmp, trim(hist))
      ier = mhd_write_dump(trim(fname)//c_null_char, dmp, fp, iprec)
      if (ier /= 0) call die(13_c_int, 'writing '//trim(fname))
      print *, ' wrote ', trim(fname), ' (', iprec, '-byte reals)'
    end if
  end do
  call system_clock(clk1)

  print *
  print *, ' steps: ', ntotal, '  wall time (s): ', real(clk1 - clk0) / real(rate)
  print *, ' Mcell-steps/s: ', real(nx) * real(ny) * real(nz) * real(ntotal) / &
                                (real(clk1 - clk0) / real(rate)) / 1.0e6
  call mhd_destroy(ctx)

contains

  subroutine die(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    print *, 'ERROR: ', what, ': ', trim(mhd_message(code)), ' (code ', code, ')'
    if (c_associated(ctx)) call mhd_destroy(ctx)
    ! the reference's abort_comms(code): stop with the code
    if (code > 0) then
      error stop code
    else
      error stop 1
    end if
  end subroutine

  ! src/initial.F90:142-195 (flag 2: Ra and Q given, CK and F derived)
  subroutine read_input(path, flag, nframe, ntotal, ndu06, ndu16, ndurestart, nx, ny, nz)
    character(len=*), intent(in) :: path
    integer, intent(out) :: flag, nframe, ntotal, ndu06, ndu16, ndurestart, nx, ny, nz
    integer :: ios
    open(3, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) then
      print *, 'Unable to open input file ', path
      error stop 1
    end if
    read(3,*)
    read(3,*)
    read(3,*) flag
    read(3,*)
    read(3,*)
    read(3,*) prm%gamma
    read(3,*) prm%sigma
    read(3,*) prm%theta
    read(3,*) prm%ck
    read(3,*) prm%cm
    read(3,*) prm%f
    read(3,*) prm%tayl
    read(3,*) prm%psi
    read(3,*) prm%zeta
    read(3,*) prm%xmax
    read(3,*) prm%ymax
    read(3,*) prm%sf
    read(3,*) nframe
    read(3,*) ntotal
    read(3,*) ndu06
    read(3,*) ndu16
    read(3,*) ndurestart
    read(3,*) nx
    read(3,*) ny
    read(3,*) nz
    close(3)
    if (flag == 2) then
      rayl = prm%ck
      chand = prm%f
      call mhd_ra_q_to_ck_f(prm%gamma, prm%sigma, prm%theta, prm%cm, prm%zeta, rayl, chand, &
                            prm%ck, prm%f)
    else if (flag /= 1) then
      print *, 'Unknown input flag ', flag
      error stop 1
    end if
    if (nframe < 0 .or. nframe > 2) then
      print *, 'Unknown value for nframe ', nframe
      error stop 1
    end if
  end subroutine

  subroutine read_extensions(path)
    character(len=*), intent(in) :: path
    character(len=32) :: key, val
    integer :: ios
    logical :: there
    inquire(file=path, exist=there)
    if (.not. there) return
    open(4, file=path, status='old', action='read')
    do
      read(4, *, iostat=ios) key, val
      if (ios /= 0) exit
      select case (trim(key))
      case ('dt_fixed')
        read(val, *) dt_fixed
      case ('t_amp')
        read(val, *) prm%t_amp
      case ('b_amp')
        read(val, *) prm%b_amp
      case ('precision')
        if (trim(val) == 'mixed') prec = MHD_PREC_MIXED
      case ('eta_c')
        read(val, *) prm%eta_c
      case ('eta_z0')
        read(val, *) prm%eta_z0
      case ('eta_w')
        read(val, *) prm%eta_w
      case ('radbc')
        if (trim(val) == '1') prm%opts = ior(prm%opts, MHD_OPT_RADBC)
      case ('constflux')
        if (trim(val) == '1') prm%opts = ior(prm%opts, MHD_OPT_CONSTFLUX)
      case ('perfect')
        if (trim(val) == '1') prm%opts = ior(prm%opts, MHD_OPT_PERFECT)
      case ('kinematic')
        if (trim(val) == '1') prm%opts = ior(prm%opts, MHD_OPT_KINEMATIC)
      case ('dynamo')
        if (trim(val) == '1') prm%opts = ior(prm%opts, MHD_OPT_DYNAMO)
      end select
    end do
    close(4)
  end subroutine

  subroutine set_history(d, text)
    type(mhd_dump_t), intent(inout) :: d
    character(len=*), intent(in) :: text
    integer :: i
    d%history = c_null_char
    do i = 1, min(len(text), 255)
      d%history(i) = text(i:i)
    end do
  end subroutine

end program p3d_gpu
