!> A DIM field resident in HBM, driven from Fortran through the C ABI.
!>
!> Mirrors the calls the reference solver makes on its field for the hot path
!> (jdesmarais/dim2D_finite_volume):
!>   field%ini coordinates / borders        field/field_abstract_class.f:244-315
!>   field%integrate(dt)                    field/field_class.f:61-72
!>     -> td_integrator%integrate           td_integrator/rungekutta3rdtvd/td_integrator_class.f:103-161
!>   td_operators%compute_time_dev          td_operators/finitevolume/td_operators_class.f:119-235
!>   bc_operators_gen%apply_bc_on_nodes     boundary_conditions/bc_operators_gen_class.f:238-381
!>   field%check_steady_state              field/field_abstract_class.f:629-667
!>   field_extended%adapt_domain          field/field_extended_class.f:427-452
!>     (rectangular growth of open edges; the grid, maps and mirror follow)
!> The host array nodes(nx,ny,ne) is a mirror: it is refreshed by sync_nodes
!> (before output, as field_abstract%write_data would need) and pushed by
!> set_nodes; integrate keeps it current after every step when
!> sync_each_step is set (the reference's behaviour), integrate_steps does not.
module dim2d_hip_field

  use, intrinsic :: iso_c_binding
  use dim2d_hip_cabi

  implicit none

  private
  public :: hip_field

  integer, parameter :: rkind = c_double
  integer, parameter :: ne = 4
  integer, parameter :: bc_size = 2

  type :: hip_field
    type(c_ptr) :: ctx = c_null_ptr
    type(dim2d_params) :: params
    integer :: nx = 0, ny = 0
    real(rkind), allocatable :: nodes(:,:,:)
    real(rkind), allocatable :: x_map(:), y_map(:)
    integer, dimension(2) :: x_borders, y_borders
    real(rkind) :: dx = 0, dy = 0, time = 0
    logical :: sync_each_step = .true.
  contains
    procedure :: ini
    procedure :: set_nodes
    procedure :: sync_nodes
    procedure :: integrate
    procedure :: integrate_steps
    procedure :: compute_time_dev
    procedure :: apply_bc_on_nodes
    procedure :: cfl_dt
    procedure :: diagnostics
    procedure :: check_steady_state
    procedure :: adapt_domain
    procedure :: finalize
  end type hip_field

contains

  !> coordinates (ini_coordinates), borders (ini_for_timeInt) and the device context
  subroutine ini(this, params)
    class(hip_field), intent(inout) :: this
    type(dim2d_params), intent(in) :: params
    call set_grid(this, params)
    this%nodes = 0
    call dim2d_check(dim2d_ctx_create(this%params, this%ctx), 'dim2d_ctx_create')
  end subroutine ini

  !> grid size, coordinates and borders of `params`; the host mirror is
  !> (re)allocated
  subroutine set_grid(this, params)
    class(hip_field), intent(inout) :: this
    type(dim2d_params), intent(in) :: params
    integer :: i, j
    this%params = params
    this%nx = params%nx
    this%ny = params%ny
    if (allocated(this%nodes)) deallocate(this%nodes)
    if (allocated(this%x_map)) deallocate(this%x_map)
    if (allocated(this%y_map)) deallocate(this%y_map)
    allocate(this%nodes(this%nx, this%ny, ne))
    allocate(this%x_map(this%nx), this%y_map(this%ny))
    this%dx = (params%x_max - params%x_min) / (this%nx - 1 - 2*bc_size)
    this%dy = (params%y_max - params%y_min) / (this%ny - 1 - 2*bc_size)
    do i = 1, this%nx
      this%x_map(i) = params%x_min + (i - 1 - bc_size)*this%dx
    end do
    do j = 1, this%ny
      this%y_map(j) = params%y_min + (j - 1 - bc_size)*this%dy
    end do
    this%x_borders = [bc_size + 1, this%nx - bc_size]
    this%y_borders = [bc_size + 1, this%ny - bc_size]
  end subroutine set_grid

  subroutine set_nodes(this, nodes)
    class(hip_field), intent(inout) :: this
    real(rkind), intent(in) :: nodes(:,:,:)
    this%nodes = nodes
    call dim2d_check(dim2d_upload(this%ctx, this%nodes), 'dim2d_upload')
  end subroutine set_nodes

  subroutine sync_nodes(this)
    class(hip_field), intent(inout) :: this
    call dim2d_check(dim2d_download(this%ctx, this%nodes), 'dim2d_download')
  end subroutine sync_nodes

  !> field%integrate(dt): one RK3-TVD step, then time = time + dt
  subroutine integrate(this, dt)
    class(hip_field), intent(inout) :: this
    real(rkind), intent(in) :: dt
    call dim2d_check(dim2d_rk3_step(this%ctx, this%time, dt), 'dim2d_rk3_step')
    this%time = this%time + dt
    if (this%sync_each_step) call this%sync_nodes()
  end subroutine integrate

  !> nsteps steps on the device (fixed dt, or the CFL rule when cfl > 0)
  subroutine integrate_steps(this, nsteps, dt, cfl)
    class(hip_field), intent(inout) :: this
    integer, intent(in) :: nsteps
    real(rkind), intent(in) :: dt, cfl
    real(c_double) :: tt(2)
    call dim2d_check(dim2d_integrate(this%ctx, this%time, dt, cfl, int(nsteps, c_int32_t)), &
                     'dim2d_integrate')
    call dim2d_check(dim2d_time(this%ctx, tt), 'dim2d_time')
    this%time = tt(1)
  end subroutine integrate_steps

  !> td_operators%compute_time_dev of the device state
  function compute_time_dev(this) result(time_dev)
    class(hip_field), intent(inout) :: this
    real(rkind), dimension(this%nx, this%ny, ne) :: time_dev
    call dim2d_check(dim2d_time_dev(this%ctx, this%time, time_dev), 'dim2d_time_dev')
  end function compute_time_dev

  subroutine apply_bc_on_nodes(this)
    class(hip_field), intent(inout) :: this
    call dim2d_check(dim2d_apply_bc(this%ctx, this%time), 'dim2d_apply_bc')
  end subroutine apply_bc_on_nodes

  function cfl_dt(this, cfl) result(dt)
    class(hip_field), intent(inout) :: this
    real(rkind), intent(in) :: cfl
    real(rkind) :: dt
    real(c_double) :: out(4)
    call dim2d_check(dim2d_cfl_dt(this%ctx, cfl, out), 'dim2d_cfl_dt')
    dt = out(1)
  end function cfl_dt

  function diagnostics(this) result(out)
    class(hip_field), intent(inout) :: this
    real(rkind) :: out(7)
    call dim2d_check(dim2d_diagnostics(this%ctx, out), 'dim2d_diagnostics')
  end function diagnostics

  !> field%check_steady_state (field_abstract_class.f:629-667): max |time_dev|
  !> over the cells with a derivative below steady_state_limit
  function check_steady_state(this, limit) result(steady)
    class(hip_field), intent(inout) :: this
    real(rkind), intent(in) :: limit
    logical :: steady
    real(c_double) :: out(4)
    integer(c_int32_t) :: st
    call dim2d_check(dim2d_check_steady_state(this%ctx, limit, out, st), &
                     'dim2d_check_steady_state')
    steady = st /= 0
  end function check_steady_state

  !> field_extended%adapt_domain: each open edge in adapt_mask bits 4..7 (16 W,
  !> 32 E, 64 S, 128 N) that has grown and whose decrease band holds no
  !> undermined point goes back to its place at creation; then each open edge
  !> in bits 0..3 (1 W, 2 E, 4 S, 8 N) whose band of `distance` points is
  !> undermined moves out by `grow` points (md_threshold < 0: imaginary sound
  !> speed only); the host grid, maps and mirror follow, and the mirror is
  !> refreshed when the grid changed
  function adapt_domain(this, adapt_mask, distance, md_threshold, grow) result(added)
    class(hip_field), intent(inout) :: this
    integer, intent(in) :: adapt_mask, distance, grow
    real(rkind), intent(in) :: md_threshold
    integer(c_int32_t) :: added(4)
    type(dim2d_params) :: q
    call dim2d_check(dim2d_adapt_domain(this%ctx, int(adapt_mask, c_int32_t), &
                     int(distance, c_int32_t), md_threshold, int(grow, c_int32_t), added), &
                     'dim2d_adapt_domain')
    if (any(added /= 0)) then
      call dim2d_check(dim2d_get_params(this%ctx, q), 'dim2d_get_params')
      call set_grid(this, q)
      call this%sync_nodes()
    end if
  end function adapt_domain

  subroutine finalize(this)
    class(hip_field), intent(inout) :: this
    integer(c_int) :: st
    if (c_associated(this%ctx)) then
      st = dim2d_ctx_destroy(this%ctx)
      this%ctx = c_null_ptr
    end if
  end subroutine finalize

end module dim2d_hip_field
