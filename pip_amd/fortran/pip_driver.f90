! pip_driver.f90 -- Fortran driver of the two-fluid step through pipgpu.
!
! Runs the time loop of main.f90:31-68 with the procedures replaced by the
! C-ABI: set_IR -> pip_set_ir, cfl -> pip_cfl, run_solver -> pip_run_solver,
! stop_sim -> pip_status.  The initial state is the 1-D two-fluid Brio-Wu
! tube (rho=1,p=1,By=1 | rho=0.125,p=0.1,By=-1, Bx=0.75, xi_n=0.5,
! alpha_c=100, outflow, t_end=0.1) set up as shock_tube.f90:27-195 does,
! on nx cells given as the first argument.  The final U_h and U_m are
! written as raw doubles to the file named by the second argument.
! A third argument "fused" runs the same loop with pip_advance instead.
program pip_driver
  use iso_c_binding
  use pipgpu
  implicit none
  integer :: nx, ix, jx, i, nvar_m, nt, ios
  integer(c_int) :: flags, ndone
  double precision, allocatable :: U_h(:,:,:,:), U_m(:,:,:,:)
  double precision :: gm, xi_n, f_n, f_p, f_p_n, f_p_p, x, dx, rho, p, by
  double precision :: dt, dtnew, time, tend, tdone
  character(len=256) :: arg, outfile
  logical :: fused
  type(pip_cfg) :: cfg
  type(c_ptr) :: ctx

  nx = 1024
  outfile = 'pip_driver_out.bin'
  fused = .false.
  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg)
    read(arg, *, iostat=ios) nx
  endif
  if (command_argument_count() >= 2) call get_command_argument(2, outfile)
  if (command_argument_count() >= 3) then
    call get_command_argument(3, arg)
    fused = (trim(arg) == 'fused')
  endif

  ix = nx + 4
  jx = 1
  nvar_m = 8
  gm = 5.0d0 / 3.0d0
  xi_n = 0.5d0
  tend = 0.1d0
  dx = 1.0d0 / dble(nx)
  allocate(U_h(ix, jx, 1, 5), U_m(ix, jx, 1, nvar_m))

  ! ionisation split of shock_tube.f90:34-37
  f_n = xi_n
  f_p = 1.0d0 - xi_n
  f_p_n = f_n / (f_n + 2.0d0 * f_p)
  f_p_p = 2.0d0 * f_p / (f_n + 2.0d0 * f_p)
  do i = 1, ix
    x = 0.0d0 + (dble(i - 3) + 0.5d0) * dx
    if (x < 0.5d0) then
      rho = 1.0d0; p = 1.0d0; by = 1.0d0
    else
      rho = 0.125d0; p = 0.1d0; by = -1.0d0
    endif
    ! pv2cq_hd / pv2cq_mhd (Scheme_rot.f90:30-79) with v = 0
    U_h(i, 1, 1, :) = (/ f_n * rho, 0.0d0, 0.0d0, 0.0d0, f_p_n * p / (gm - 1.0d0) /)
    U_m(i, 1, 1, :) = (/ f_p * rho, 0.0d0, 0.0d0, 0.0d0, &
         (f_p_p * p / (gm - 1.0d0) + 0.0d0) + 0.5d0 * ((0.75d0**2 + by**2) + 0.0d0), &
         0.75d0, by, 0.0d0 /)
  enddo

  cfg%ndim = 1; cfg%ix = ix; cfg%jx = jx; cfg%nvar_m = nvar_m
  cfg%flag_divb = 0; cfg%t_order = 2; cfg%flag_ir = 0; cfg%flag_ir_type = 0
  cfg%flag_col = 1; cfg%ac_field = 0; cfg%flag_bnd = (/ 10, 10, 10, 10 /)
  cfg%gm = gm; cfg%safety = 0.3d0; cfg%ro_lim = 1.0d-8; cfg%pr_lim = 1.0d-8
  cfg%dx = dx; cfg%dy = 100.0d0; cfg%col = 100.0d0; cfg%db_clean = 0.7d0
  cfg%t_ir = 1.0d-5; cfg%T0 = 5500.0d0; cfg%n0 = 7.5d16; cfg%beta = 0.1d0
  cfg%flag_pip_imp = 0; cfg%pip_imp_factor = 10.0d0

  call pip_check(pip_create(cfg, ctx), 'pip_create')
  call pip_check(pip_upload(ctx, U_h, U_m, c_null_ptr), 'pip_upload')
  call pip_check(pip_prologue(ctx, dt), 'prologue')
  time = 0.0d0
  nt = 0
  if (fused) then
    call pip_check(pip_advance(ctx, 100000000_c_int, tend, ndone, tdone), 'pip_advance')
    nt = ndone
    time = tdone
  else
    do while (time < tend)
      if (cfg%flag_ir >= 1) call pip_check(pip_set_ir(ctx), 'set_IR')
      call pip_check(pip_cfl(ctx, dt, 1_c_int), 'cfl')
      call pip_check(pip_run_solver(ctx), 'run_solver')
      if (cfg%flag_ir >= 1) call pip_check(pip_set_ir(ctx), 'set_IR')
      call pip_check(pip_cfl(ctx, dtnew, 0_c_int), 'cfl')
      time = time + dt
      nt = nt + 1
      call pip_check(pip_status(ctx, flags), 'stop_sim')
      if (flags /= 0) exit
    enddo
  endif
  call pip_check(pip_download(ctx, U_h, U_m), 'pip_download')
  call pip_check(pip_destroy(ctx), 'pip_destroy')
  open(21, file=trim(outfile), access='stream', form='unformatted', status='replace')
  write(21) U_h, U_m
  close(21)
  print '(a,i6,a,es24.16)', ' pip_driver: steps=', nt, ' time=', time
end program pip_driver
