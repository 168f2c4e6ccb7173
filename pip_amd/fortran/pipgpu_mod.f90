! pipgpu_mod.f90 -- iso_c_binding interface to libpipgpu.so (include/pipgpu.h).
!
! This is the reference-side binding: a Fortran driver of PIP (main.f90)
! uses this module instead of solver_rot/scheme_rot/pip_rot for the
! two-fluid step.  Arrays are passed as the explicit-shape
! U(ix,jx,kx,nvar) arrays of module globalvar (Globalvar.f90:17), by
! reference, with no copies.
module pipgpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: PIP_OK = 0
  integer(c_int), parameter :: PIP_STOP_DT_SMALL = 1, PIP_STOP_NAN = 2
  integer(c_int), parameter :: PIP_ARITH_EXACT = 0, PIP_ARITH_RELAXED = 1

  ! pip_cfg of include/pipgpu.h
  type, bind(c) :: pip_cfg
    integer(c_int) :: ndim, ix, jx, nvar_m, flag_divb, t_order
    integer(c_int) :: flag_ir, flag_ir_type, flag_col, ac_field
    integer(c_int) :: flag_bnd(4)
    real(c_double) :: gm, safety, ro_lim, pr_lim, dx, dy
    real(c_double) :: col, db_clean, t_ir, T0, n0, beta
    integer(c_int) :: flag_pip_imp
    real(c_double) :: pip_imp_factor
  end type pip_cfg

  interface
    integer(c_int) function pip_create(cfg, ctx) bind(c, name="pip_create")
      import :: c_int, c_ptr, pip_cfg
      type(pip_cfg), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function pip_destroy(ctx) bind(c, name="pip_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function pip_upload(ctx, U_h, U_m, ac) bind(c, name="pip_upload")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(in) :: U_h, U_m
      type(c_ptr), value :: ac
    end function
    integer(c_int) function pip_download(ctx, U_h, U_m) bind(c, name="pip_download")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), dimension(*), intent(out) :: U_h, U_m
    end function
    integer(c_int) function pip_prologue(ctx, dt) bind(c, name="pip_prologue")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dt
    end function
    integer(c_int) function pip_set_ir(ctx) bind(c, name="pip_set_ir")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function pip_cfl(ctx, dt, into_global) bind(c, name="pip_cfl")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dt
      integer(c_int), value :: into_global
    end function
    integer(c_int) function pip_run_solver(ctx) bind(c, name="pip_run_solver")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function pip_bnd(ctx) bind(c, name="pip_bnd")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function pip_advance(ctx, max_steps, tend, steps_done, time) bind(c, name="pip_advance")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: max_steps
      real(c_double), value :: tend
      integer(c_int), intent(out) :: steps_done
      real(c_double), intent(out) :: time
    end function
    integer(c_int) function pip_status(ctx, flags) bind(c, name="pip_status")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: flags
    end function
    type(c_ptr) function pip_last_error() bind(c, name="pip_last_error")
      import :: c_ptr
    end function
    ! Tuning and variants (no reference counterpart): the stage tile, the
    ! arithmetic (PIP_ARITH_EXACT, bitwise the reference path, or
    ! PIP_ARITH_RELAXED) and the hipGraph replay of pip_advance's chunks.
    integer(c_int) function pip_set_tile(ctx, tx, ty) bind(c, name="pip_set_tile")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: tx, ty
    end function
    integer(c_int) function pip_set_arith(ctx, mode) bind(c, name="pip_set_arith")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
    end function
    integer(c_int) function pip_set_step_graph(ctx, on) bind(c, name="pip_set_step_graph")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function
  end interface

contains

  ! Abort with the library's message (the reference stops on fatal
  ! configurations, e.g. Solver_rot.f90:137-138).
  subroutine pip_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    character(len=256) :: text
    integer :: k
    if (rc == PIP_OK) return
    call c_f_pointer(pip_last_error(), msg, [256])
    text = ' '
    do k = 1, 256
      if (msg(k) == c_null_char) exit
      text(k:k) = msg(k)
    end do
    print *, 'pipgpu error in ', where, ': ', trim(text)
    stop 1
  end subroutine pip_check

end module pipgpu
