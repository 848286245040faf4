! routine_driver.f90 -- predictor-corrector steps called routine by routine, in
! the order of time_iteration.f90:509-657 and single_step_compute.f90:80-146:
! Distribute_Period_Boundary_Particle, Devide_block_and_send_tasks, then per
! evaluation find_particle_pair, Define_freesurface_and_Normailized_derivates and
! the solver, the predictor and corrector stage updates, and after the step
! Reinitialize_particle_density (every mls_every steps), Shift_particle_position
! and Refresh_domain.
! Usage: routine_driver n_side nsteps solver sps_on prefix. Writes (stream access):
!   prefix_pairs.bin  int64 pair_number, pair_i(:), pair_j(:) of step 1's first evaluation
!   prefix_rates.bin  drhodt(n), particle_acceleration(n,2) of that evaluation
!   prefix_state.bin  position(n,2), velocity(n,2), rho(n), press(n) after nsteps
program routine_driver
  use iso_c_binding
  use sph2d_interface
  implicit none
  type(sph2d_params) :: prm
  type(c_ptr) :: ctx
  integer(c_int) :: rc
  integer :: n_side, nsteps, it, n, u
  integer(c_int32_t) :: variant, sps_on
  integer(c_int64_t) :: np
  character(len=256) :: arg, prefix
  integer(c_int32_t), allocatable :: pair_i(:), pair_j(:), ptype(:)
  integer(c_int32_t) :: dummy_i(1), dummy_j(1)
  real(c_double), allocatable :: drhodt(:), acc(:, :), pos(:, :), vel(:, :), rho(:), press(:), c(:), mass(:)

  call get_command_argument(1, arg); read(arg, *) n_side
  call get_command_argument(2, arg); read(arg, *) nsteps
  call get_command_argument(3, arg); read(arg, *) variant
  call get_command_argument(4, arg); read(arg, *) sps_on
  call get_command_argument(5, prefix)

  rc = sph2d_default_params(prm); call check(rc)
  prm%nx = n_side
  prm%ny = n_side
  prm%dx = 1.0d0 / n_side
  prm%h = 1.20d0 * prm%dx
  prm%max_shift = 0.25d0 * prm%dx
  prm%dt = 1.0d-4 * (100.0d0 / n_side)
  prm%integrator = SPH2D_INTEGRATOR_PC
  prm%solver = variant
  prm%sps_on = sps_on
  n = n_side * n_side
  rc = sph2d_ctx_create(prm, n, ctx); call check(rc)
  rc = sph2d_init_tgv(ctx); call check(rc)

  do it = 1, nsteps
    rc = sph2d_distribute_period_boundary(ctx); call check(rc)
    rc = sph2d_devide_block(ctx, it); call check(rc)
    if (it == 1) then
      call evaluate(it)
      rc = sph2d_download_pair_list(ctx, 0_c_int64_t, dummy_i, dummy_j, np); call check(rc)
      allocate(pair_i(np), pair_j(np))
      rc = sph2d_download_pair_list(ctx, np, pair_i, pair_j, np); call check(rc)
      allocate(drhodt(n), acc(n, 2))
      rc = sph2d_download_rates(ctx, n, drhodt, acc, n); call check(rc)
      open(newunit=u, file=trim(prefix)//'_pairs.bin', access='stream', form='unformatted', status='replace')
      write(u) np, pair_i, pair_j
      close(u)
      open(newunit=u, file=trim(prefix)//'_rates.bin', access='stream', form='unformatted', status='replace')
      write(u) drhodt, acc
      close(u)
    end if
    rc = sph2d_stage_update(ctx, SPH2D_INTEGRATOR_PC, 1); call check(rc)        ! predictor (:553-576)
    call evaluate(it)
    rc = sph2d_stage_update(ctx, SPH2D_INTEGRATOR_PC, 2); call check(rc)        ! corrector (:590-599)
    if (mod(it, prm%mls_every) == 0) then
      rc = sph2d_reinitialize_particle_density(ctx, it); call check(rc)
    end if
    rc = sph2d_shift_particle_position(ctx, it); call check(rc)
    rc = sph2d_refresh_domain(ctx, it); call check(rc)
  end do

  allocate(pos(n, 2), vel(n, 2), rho(n), press(n), c(n), mass(n), ptype(n))
  rc = sph2d_download_state(ctx, n, pos, vel, rho, press, c, mass, ptype, n); call check(rc)
  open(newunit=u, file=trim(prefix)//'_state.bin', access='stream', form='unformatted', status='replace')
  write(u) pos, vel, rho, press
  close(u)
  rc = sph2d_ctx_destroy(ctx)
  write(*, '(a)') 'ok'

contains

  ! single_step_compute.f90:80-146 routine by routine
  subroutine evaluate(step)
    integer, intent(in) :: step
    rc = sph2d_find_particle_pair(ctx, step); call check(rc)
    rc = sph2d_define_freesurface(ctx); call check(rc)
    rc = sph2d_solver(ctx, variant, sps_on); call check(rc)
  end subroutine evaluate

  subroutine check(code)
    integer(c_int), intent(in) :: code
    character(kind=c_char), pointer :: msg(:)
    integer :: k
    if (code /= 0) then
      call c_f_pointer(sph2d_last_error(), msg, (/ 256 /))
      write(*, '(a, i3)', advance='no') 'sph2d error ', code
      do k = 1, 256
        if (msg(k) == c_null_char) exit
        write(*, '(a)', advance='no') msg(k)
      end do
      write(*, *)
      stop 1
    end if
  end subroutine check
end program routine_driver
