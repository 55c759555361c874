!> A parameter sweep of independent simulations run as one batch of the HIP
!> library (dim2d_batch_*): where the reference's sim_dim2d (src/main/sim_dim2d.f:16-115)
!> integrates one field per process, this driver integrates every simulation
!> of a list in the same persistent launches, each with its own grid, edges
!> and fixed time step.  No counterpart in the reference.
!>
!>   sweep_dim2d_hip list.txt in.bin out.bin nt precision [fuse [tune_steps]]
!>
!> list.txt: one simulation per line, "nx ny bc_w bc_e bc_s bc_n dt".
!> in.bin / out.bin: the simulations' raw float64 nodes(nx,ny,4) in Fortran
!> order, one after the other in list order.  fuse: -1 (default) keeps the
!> library's default, 0 runs the three stages apart, 1 fuses stages 1 and 2,
!> 2 fuses and pipelines.  tune_steps > 0 first runs the library's tile
!> autotuner on that many steps (the state is left as it was).  Test physics
!> of the reference (dim2d_parameters.f:182-187) on [-0.5, 0.5]^2.
program sweep_dim2d_hip

  use, intrinsic :: iso_c_binding
  use dim2d_hip_cabi

  implicit none

  character(len=512) :: flist, fin, fout, arg
  type(dim2d_params), allocatable :: p(:)
  real(c_double), allocatable :: dts(:), nodes(:)
  integer :: n, i, nt, prec, fuse, tune, u, ios, nx, ny, bw, be, bs, bn
  integer(c_int32_t) :: tiles(6)
  real(c_double) :: dt, tuned(3)
  type(c_ptr) :: b
  integer(8) :: c0, c1, crate, cells

  if (command_argument_count() < 5) then
    print '(A)', 'usage: sweep_dim2d_hip list.txt in.bin out.bin nt precision [fuse [tune_steps]]'
    stop 2
  end if
  call get_command_argument(1, flist)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)
  call int_arg(4, nt)
  call int_arg(5, prec)
  fuse = -1
  tune = 0
  if (command_argument_count() >= 6) call int_arg(6, fuse)
  if (command_argument_count() >= 7) call int_arg(7, tune)

  ! the list: one simulation per line
  open(newunit=u, file=trim(flist), status='old', action='read')
  n = 0
  do
    read(u, *, iostat=ios) nx, ny, bw, be, bs, bn, dt
    if (ios /= 0) exit
    n = n + 1
  end do
  if (n == 0) then
    print '(A)', 'sweep_dim2d_hip: empty list'
    stop 2
  end if
  rewind(u)
  allocate(p(n), dts(n))
  cells = 0
  do i = 1, n
    read(u, *) nx, ny, bw, be, bs, bn, dts(i)
    p(i)%nx = nx; p(i)%ny = ny; p(i)%bc_size = 2
    p(i)%bc_w = bw; p(i)%bc_e = be; p(i)%bc_s = bs; p(i)%bc_n = bn
    p(i)%precision = prec; p(i)%gravity_ac = 0
    p(i)%x_min = -0.5d0; p(i)%x_max = 0.5d0; p(i)%y_min = -0.5d0; p(i)%y_max = 0.5d0
    p(i)%re = 5.0d0; p(i)%we = 10.0d0; p(i)%pr = 20.0d0
    p(i)%viscous_r = -1.5d0; p(i)%cv_r = 2.5d0
    p(i)%gravity = 0.0d0; p(i)%wall_contact_angle = 90.0d0; p(i)%wall_heat_flux = 0.0d0
    p(i)%integrator = DIM2D_INT_RK3TVD; p(i)%operators = DIM2D_OPS_MATTSSON
    cells = cells + int(nx - 4, 8) * int(ny - 4, 8)
  end do
  close(u)

  call dim2d_check(dim2d_batch_create(p, int(n, c_int32_t), b), 'dim2d_batch_create')
  if (fuse >= 0) call dim2d_check(dim2d_batch_set_stage_fusion(b, int(fuse, c_int32_t), 0, 0), &
                                  'dim2d_batch_set_stage_fusion')

  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', &
       action='read')
  do i = 1, n
    allocate(nodes(p(i)%nx * p(i)%ny * 4))
    read(u) nodes
    call dim2d_check(dim2d_batch_upload(b, int(i - 1, c_int32_t), nodes), 'dim2d_batch_upload')
    deallocate(nodes)
  end do
  close(u)

  if (tune > 0) call dim2d_check(dim2d_batch_autotune(b, dts, int(tune, c_int32_t), tuned), &
                                 'dim2d_batch_autotune')
  call dim2d_check(dim2d_batch_get_tiles(b, tiles), 'dim2d_batch_get_tiles')
  print '(A,6I8)', 'tiles:', tiles

  call system_clock(c0, crate)
  call dim2d_check(dim2d_batch_integrate(b, dts, int(nt, c_int32_t)), 'dim2d_batch_integrate')
  call dim2d_check(dim2d_batch_synchronize(b), 'dim2d_batch_synchronize')
  call system_clock(c1)

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', &
       action='write')
  do i = 1, n
    allocate(nodes(p(i)%nx * p(i)%ny * 4))
    call dim2d_check(dim2d_batch_download(b, int(i - 1, c_int32_t), nodes), &
                     'dim2d_batch_download')
    write(u) nodes
    deallocate(nodes)
  end do
  close(u)
  call dim2d_check(dim2d_batch_destroy(b), 'dim2d_batch_destroy')

  if (c1 > c0 .and. nt > 0) then
    print '(A,I0,A,I0,A,F10.4,A,ES11.4,A)', 'simulations: ', n, ', steps: ', nt, ', ', &
      1.0d3 * dble(c1 - c0) / dble(crate) / nt, ' ms per step, ', &
      dble(cells) * nt / (dble(c1 - c0) / dble(crate)), ' updates/s'
  end if

contains

  subroutine int_arg(k, v)
    integer, intent(in) :: k
    integer, intent(out) :: v
    call get_command_argument(k, arg)
    read(arg, *) v
  end subroutine int_arg

end program sweep_dim2d_hip
