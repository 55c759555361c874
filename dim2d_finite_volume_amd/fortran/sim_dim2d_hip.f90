!> Time loop of the reference's sim_dim2d (src/main/sim_dim2d.f:16-115) on the
!> HBM-resident field: read an initial state, integrate nt steps, write the
!> final state, print the timing.
!>
!>   sim_dim2d_hip in.bin out.bin nx ny bc_w bc_e bc_s bc_n nt dt cfl precision [nsync
!>                 [adapt_every adapt_mask grow md_threshold]]
!>
!> in.bin / out.bin: raw float64 nodes(nx,ny,4) in Fortran order.  With
!> cfl > 0 the time step comes from the library's CFL rule; otherwise dt is
!> fixed.  nsync > 0 calls field%integrate step by step with the host mirror
!> refreshed every step (the reference's calling pattern); nsync = 0 (default)
!> runs the steps on the device in one call.  adapt_every > 0 (open edges)
!> runs the steps in chunks of adapt_every, each followed by
!> field%adapt_domain(adapt_mask, 5, md_threshold, grow), the reference's
!> field_extended loop (integrate_ext then adapt_domain); out.bin then has the
!> final grid, printed as "grid: nx ny".  Test physics of the reference
!> (dim2d_parameters.f:182-187) on [-0.5, 0.5]^2.
program sim_dim2d_hip

  use, intrinsic :: iso_c_binding
  use dim2d_hip_cabi
  use dim2d_hip_field

  implicit none

  type(hip_field) :: f
  type(dim2d_params) :: p
  character(len=512) :: fin, fout, arg
  integer :: nx, ny, bw, be, bs, bn, nt, prec, nsync, u, t
  integer :: adapt_every, adapt_mask, grow, done, chunk
  integer(c_int32_t) :: added(4)
  real(c_double) :: dt, cfl, diag(7), md_threshold
  real :: time1, time2
  integer(8) :: c0, c1, crate

  if (command_argument_count() < 12) then
    print '(A)', 'usage: sim_dim2d_hip in.bin out.bin nx ny bc_w bc_e bc_s bc_n nt dt cfl precision [nsync]'
    stop 2
  end if
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call int_arg(3, nx); call int_arg(4, ny)
  call int_arg(5, bw); call int_arg(6, be); call int_arg(7, bs); call int_arg(8, bn)
  call int_arg(9, nt)
  call real_arg(10, dt); call real_arg(11, cfl)
  call int_arg(12, prec)
  nsync = 0
  if (command_argument_count() >= 13) call int_arg(13, nsync)
  adapt_every = 0; adapt_mask = 0; grow = 0; md_threshold = -1.0d0
  if (command_argument_count() >= 17) then
    call int_arg(14, adapt_every); call int_arg(15, adapt_mask); call int_arg(16, grow)
    call real_arg(17, md_threshold)
  end if

  p%nx = nx; p%ny = ny; p%bc_size = 2
  p%bc_w = bw; p%bc_e = be; p%bc_s = bs; p%bc_n = bn
  p%precision = prec; p%gravity_ac = 0
  p%x_min = -0.5d0; p%x_max = 0.5d0; p%y_min = -0.5d0; p%y_max = 0.5d0
  p%re = 5.0d0; p%we = 10.0d0; p%pr = 20.0d0; p%viscous_r = -1.5d0; p%cv_r = 2.5d0
  p%gravity = 0.0d0; p%wall_contact_angle = 90.0d0; p%wall_heat_flux = 0.0d0

  call f%ini(p)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) f%nodes
  close(u)
  call f%set_nodes(f%nodes)

  call cpu_time(time1)
  call system_clock(c0, crate)
  if (adapt_every > 0) then
    done = 0
    do while (done < nt)
      chunk = min(adapt_every, nt - done)
      call f%integrate_steps(chunk, dt, cfl)
      done = done + chunk
      added = f%adapt_domain(adapt_mask, 5, md_threshold, grow)
    end do
    call f%sync_nodes()
  else if (nsync > 0) then
    f%sync_each_step = .true.
    do t = 1, nt
      if (cfl > 0) dt = f%cfl_dt(cfl)
      call f%integrate(dt)
    end do
  else
    call f%integrate_steps(nt, dt, cfl)
    call f%sync_nodes()
  end if
  call system_clock(c1)
  call cpu_time(time2)

  diag = f%diagnostics()
  print '(A,I0,A,I0)', 'grid: ', f%nx, ' ', f%ny
  print '(A,I0,A,ES12.5,A,ES12.5)', 'steps: ', nt, '  time: ', f%time, '  mass: ', diag(1)
  print '(A,F10.4,A,F10.4,A)', 'cpu time: ', time2 - time1, ' s, wall time: ', &
    real(c1 - c0)/real(crate), ' s'

  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(u) f%nodes
  close(u)
  call f%finalize()

contains

  subroutine int_arg(k, v)
    integer, intent(in) :: k
    integer, intent(out) :: v
    call get_command_argument(k, arg)
    read(arg, *) v
  end subroutine int_arg

  subroutine real_arg(k, v)
    integer, intent(in) :: k
    real(c_double), intent(out) :: v
    call get_command_argument(k, arg)
    read(arg, *) v
  end subroutine real_arg

end program sim_dim2d_hip
