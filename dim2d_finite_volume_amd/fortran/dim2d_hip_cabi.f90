!> ISO_C_BINDING interface to the HIP library's C ABI (include/dim2d_hip.h).
!>
!> Every function returns a C int status (0 = success); dim2d_last_error gives
!> the message of the last failure on the calling thread.
module dim2d_hip_cabi

  use, intrinsic :: iso_c_binding

  implicit none

  integer(c_int), parameter :: DIM2D_OK = 0
  integer(c_int), parameter :: DIM2D_BC_REFLECTION = 0
  integer(c_int), parameter :: DIM2D_BC_PERIODIC = 1
  integer(c_int), parameter :: DIM2D_BC_WALL = 2
  integer(c_int), parameter :: DIM2D_BC_OPEN = 3
  integer(c_int), parameter :: DIM2D_PREC_FP64 = 0
  integer(c_int), parameter :: DIM2D_PREC_MIXED = 1
  integer(c_int), parameter :: DIM2D_INT_RK3TVD = 0
  integer(c_int), parameter :: DIM2D_INT_LSRK3 = 1
  integer(c_int), parameter :: DIM2D_OPS_MATTSSON = 0
  integer(c_int), parameter :: DIM2D_OPS_COCKBURN_GAU = 1

  !> dim2d_params
  type, bind(C) :: dim2d_params
    integer(c_int32_t) :: nx, ny, bc_size
    integer(c_int32_t) :: bc_w, bc_e, bc_s, bc_n
    integer(c_int32_t) :: precision, gravity_ac
    real(c_double) :: x_min, x_max, y_min, y_max
    real(c_double) :: re, we, pr, viscous_r, cv_r
    real(c_double) :: gravity, wall_contact_angle, wall_heat_flux
    integer(c_int32_t) :: integrator = 0
    integer(c_int32_t) :: operators = 0
  end type dim2d_params

  interface

    function dim2d_last_error() bind(C, name="dim2d_last_error") result(msg)
      import :: c_ptr
      type(c_ptr) :: msg
    end function dim2d_last_error

    function dim2d_ctx_create(params, ctx) bind(C, name="dim2d_ctx_create") result(st)
      import :: c_int, c_ptr, dim2d_params
      type(dim2d_params), intent(in) :: params
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: st
    end function dim2d_ctx_create

    function dim2d_ctx_destroy(ctx) bind(C, name="dim2d_ctx_destroy") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: st
    end function dim2d_ctx_destroy

    function dim2d_upload(ctx, nodes) bind(C, name="dim2d_upload") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: nodes(*)
      integer(c_int) :: st
    end function dim2d_upload

    function dim2d_download(ctx, nodes) bind(C, name="dim2d_download") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: nodes(*)
      integer(c_int) :: st
    end function dim2d_download

    function dim2d_time_dev(ctx, t, time_dev) bind(C, name="dim2d_time_dev") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t
      real(c_double), intent(out) :: time_dev(*)
      integer(c_int) :: st
    end function dim2d_time_dev

    function dim2d_apply_bc(ctx, t) bind(C, name="dim2d_apply_bc") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t
      integer(c_int) :: st
    end function dim2d_apply_bc

    function dim2d_rk3_step(ctx, t, dt) bind(C, name="dim2d_rk3_step") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t, dt
      integer(c_int) :: st
    end function dim2d_rk3_step

    function dim2d_integrate(ctx, t0, dt_fixed, cfl, nsteps) bind(C, name="dim2d_integrate") result(st)
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), value :: t0, dt_fixed, cfl
      integer(c_int32_t), value :: nsteps
      integer(c_int) :: st
    end function dim2d_integrate

    function dim2d_cfl_dt(ctx, cfl, out) bind(C, name="dim2d_cfl_dt") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: cfl
      real(c_double), intent(out) :: out(4)
      integer(c_int) :: st
    end function dim2d_cfl_dt

    function dim2d_diagnostics(ctx, out) bind(C, name="dim2d_diagnostics") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(7)
      integer(c_int) :: st
    end function dim2d_diagnostics

    function dim2d_time(ctx, out) bind(C, name="dim2d_time") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(2)
      integer(c_int) :: st
    end function dim2d_time

    function dim2d_synchronize(ctx) bind(C, name="dim2d_synchronize") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: st
    end function dim2d_synchronize

    function dim2d_set_tiles(ctx, block_x, rows) bind(C, name="dim2d_set_tiles") result(st)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: block_x, rows
      integer(c_int) :: st
    end function dim2d_set_tiles

    function dim2d_check_steady_state(ctx, limit, out, steady) &
        bind(C, name="dim2d_check_steady_state") result(st)
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), value :: limit
      real(c_double), intent(out) :: out(4)
      integer(c_int32_t), intent(out) :: steady
      integer(c_int) :: st
    end function dim2d_check_steady_state

    function dim2d_get_params(ctx, out) bind(C, name="dim2d_get_params") result(st)
      import :: c_int, c_ptr, dim2d_params
      type(c_ptr), value :: ctx
      type(dim2d_params), intent(out) :: out
      integer(c_int) :: st
    end function dim2d_get_params

    function dim2d_openbc_undermined(ctx, distance, md_threshold, out) &
        bind(C, name="dim2d_openbc_undermined") result(st)
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: distance
      real(c_double), value :: md_threshold
      integer(c_int32_t), intent(out) :: out(4)
      integer(c_int) :: st
    end function dim2d_openbc_undermined

    function dim2d_set_far_field(ctx, far_field) bind(C, name="dim2d_set_far_field") result(st)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: far_field(4)
      integer(c_int) :: st
    end function dim2d_set_far_field

    function dim2d_extend(ctx, add_w, add_e, add_s, add_n) bind(C, name="dim2d_extend") result(st)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: add_w, add_e, add_s, add_n
      integer(c_int) :: st
    end function dim2d_extend

    function dim2d_shrink(ctx, del_w, del_e, del_s, del_n) bind(C, name="dim2d_shrink") result(st)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: del_w, del_e, del_s, del_n
      integer(c_int) :: st
    end function dim2d_shrink

    function dim2d_get_growth(ctx, out) bind(C, name="dim2d_get_growth") result(st)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: out(4)
      integer(c_int) :: st
    end function dim2d_get_growth

    function dim2d_adapt_domain(ctx, adapt_mask, distance, md_threshold, grow, out) &
        bind(C, name="dim2d_adapt_domain") result(st)
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: adapt_mask, distance, grow
      real(c_double), value :: md_threshold
      integer(c_int32_t), intent(out) :: out(4)
      integer(c_int) :: st
    end function dim2d_adapt_domain

    ! batches of independent simulations (no counterpart in the reference:
    ! a parameter sweep in one set of persistent launches instead of N runs)
    function dim2d_batch_create(params, n, batch) bind(C, name="dim2d_batch_create") result(st)
      import :: c_int, c_ptr, c_int32_t, dim2d_params
      type(dim2d_params), intent(in) :: params(*)
      integer(c_int32_t), value :: n
      type(c_ptr), intent(out) :: batch
      integer(c_int) :: st
    end function dim2d_batch_create

    function dim2d_batch_destroy(batch) bind(C, name="dim2d_batch_destroy") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int) :: st
    end function dim2d_batch_destroy

    !> nodes: the simulation's nodes(nx,ny,4), Fortran order
    function dim2d_batch_upload(batch, sim, nodes) bind(C, name="dim2d_batch_upload") result(st)
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: sim
      real(c_double), intent(in) :: nodes(*)
      integer(c_int) :: st
    end function dim2d_batch_upload

    function dim2d_batch_download(batch, sim, nodes) bind(C, name="dim2d_batch_download") result(st)
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: sim
      real(c_double), intent(out) :: nodes(*)
      integer(c_int) :: st
    end function dim2d_batch_download

    !> dts(n): each simulation's fixed time step
    function dim2d_batch_integrate(batch, dts, nsteps) bind(C, name="dim2d_batch_integrate") &
        result(st)
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: batch
      real(c_double), intent(in) :: dts(*)
      integer(c_int32_t), value :: nsteps
      integer(c_int) :: st
    end function dim2d_batch_integrate

    function dim2d_batch_synchronize(batch) bind(C, name="dim2d_batch_synchronize") result(st)
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int) :: st
    end function dim2d_batch_synchronize

    function dim2d_batch_set_tiles(batch, block_x, rows_per_block, workgroups) &
        bind(C, name="dim2d_batch_set_tiles") result(st)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: block_x, rows_per_block, workgroups
      integer(c_int) :: st
    end function dim2d_batch_set_tiles

    !> fuse: 0 separate stages, 1 stages 1 and 2 fused, 2 fused and pipelined
    function dim2d_batch_set_stage_fusion(batch, fuse, rows_per_block, workgroups) &
        bind(C, name="dim2d_batch_set_stage_fusion") result(st)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: batch
      integer(c_int32_t), value :: fuse, rows_per_block, workgroups
      integer(c_int) :: st
    end function dim2d_batch_set_stage_fusion

    function dim2d_batch_get_tiles(batch, out) bind(C, name="dim2d_batch_get_tiles") result(st)
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: batch
      integer(c_int32_t), intent(out) :: out(6)
      integer(c_int) :: st
    end function dim2d_batch_get_tiles

    function dim2d_batch_autotune(batch, dts, nsteps, out) bind(C, name="dim2d_batch_autotune") &
        result(st)
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: batch
      real(c_double), intent(in) :: dts(*)
      integer(c_int32_t), value :: nsteps
      real(c_double), intent(out) :: out(3)
      integer(c_int) :: st
    end function dim2d_batch_autotune

    function c_strlen(s) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function c_strlen

  end interface

contains

  !> text of the last library error
  function dim2d_error_text() result(text)
    character(len=:), allocatable :: text
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = dim2d_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: text)
    do i = 1, n
      text(i:i) = chars(i)
    end do
  end function dim2d_error_text

  !> the reference's error convention (errors_module.f:69,106): print and stop
  subroutine dim2d_check(st, where)
    integer(c_int), intent(in) :: st
    character(*), intent(in) :: where
    if (st /= DIM2D_OK) then
      print '(A)', 'dim2d_hip: '//where//' failed: '//dim2d_error_text()
      stop 1
    end if
  end subroutine dim2d_check

end module dim2d_hip_cabi
