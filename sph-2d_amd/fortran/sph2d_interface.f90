! sph2d_interface.f90 -- ISO_C_BINDING interface to libsph2d.so (include/sph2d.h).
!
! A Fortran driver of mini-walker/SPH-2D `use`s this module and calls the GPU
! routines with the reference's own argument lists: compute_kernel and
! status_equation take their arguments by reference exactly as
! Taylor_Green_Vortex/compute_kernel.f90:21-27 and status_equation.f90:19-24
! declare them, so existing call sites only change the routine name.
module sph2d_interface
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: SPH2D_INTEGRATOR_PC = 2, SPH2D_INTEGRATOR_RK = 3
  integer(c_int), parameter :: SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN = 0, SPH2D_SOLVER_LINEAR_RIEMANN = 1, &
                               SPH2D_SOLVER_HLLC = 2, SPH2D_SOLVER_MUSCL = 3
  integer(c_int), parameter :: SPH2D_DAMPING_OFF = 0, SPH2D_DAMPING_COSINE = 1, SPH2D_DAMPING_OPENFOAM = 2, &
                               SPH2D_DAMPING_DUALSPH = 3
  integer(c_int), parameter :: SPH2D_WAVE_SOLITARY = 0, SPH2D_WAVE_STOKES2 = 1, SPH2D_WAVE_AIRY = 2

  type, bind(C) :: sph2d_wavemaker
    integer(c_int32_t) :: wave_type, soft_steps, standing_steps
    real(c_double) :: dt, g, water_depth, wave_height, wave_number, wave_frequency, phase
  end type sph2d_wavemaker

  type, bind(C) :: sph2d_params
    integer(c_int32_t) :: struct_size
    integer(c_int32_t) :: nx, ny
    real(c_double) :: box_x, box_y, rho0, c0sq, nu, U, dx, h, dt
    integer(c_int32_t) :: integrator, solver, precision, rate_reuse
    integer(c_int32_t) :: shift_on, mls_on, mls_every, mls_order, sample_every, periodic
    real(c_double) :: alpha, beta, fai, delta, delta_alpha, epsilon_xsph
    real(c_double) :: shift_beta, shift_R, max_shift, mls_alpha
    integer(c_int32_t) :: damping
    real(c_double) :: damping_length, damping_amplitude, water_depth
    integer(c_int32_t) :: walls
    real(c_double) :: gravity
    integer(c_int32_t) :: wavemaker_layers
    integer(c_int32_t) :: sps_on
    integer(c_int32_t) :: renormalization
    integer(c_int32_t) :: ms
    real(c_double) :: background_p, release_time, body_cx, body_cy, body_len, body_speed, body_ramp
  end type sph2d_params

  interface
    function sph2d_default_params(p) bind(C, name="sph2d_default_params") result(rc)
      import :: c_int, sph2d_params
      type(sph2d_params), intent(out) :: p
      integer(c_int) :: rc
    end function

    function sph2d_ctx_create(p, capacity, ctx) bind(C, name="sph2d_ctx_create") result(rc)
      import :: c_int, c_int32_t, c_ptr, sph2d_params
      type(sph2d_params), intent(in) :: p
      integer(c_int32_t), value :: capacity
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: rc
    end function

    function sph2d_ctx_destroy(ctx) bind(C, name="sph2d_ctx_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function sph2d_init_tgv(ctx) bind(C, name="sph2d_init_tgv") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    ! the moving-square program (Moving_Square_Or_Cylinder_Inside_a_Rectangular_Box/)
    ! x, y, btype, init may be c_null_ptr (counts only) or c_loc of arrays of cap entries
    function sph2d_moving_square_layout(p, cap, x, y, btype, init, n_fluid, n_total) &
        bind(C, name="sph2d_moving_square_layout") result(rc)
      import :: c_int, c_int32_t, c_ptr, sph2d_params
      type(sph2d_params), intent(in) :: p
      integer(c_int32_t), value :: cap
      type(c_ptr), value :: x, y, btype, init
      integer(c_int32_t), intent(out) :: n_fluid, n_total
      integer(c_int) :: rc
    end function
    function sph2d_init_moving_square(ctx) bind(C, name="sph2d_init_moving_square") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function
    function sph2d_body_motion(ctx, t) bind(C, name="sph2d_body_motion") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t
      integer(c_int) :: rc
    end function
    function sph2d_body_force(ctx, out6, step) bind(C, name="sph2d_body_force") result(rc)
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double) :: out6(6)
      integer(c_int32_t) :: step
      integer(c_int) :: rc
    end function

    ! particle_position(n_total,dim) etc. pass straight through: (ld,2) column-major
    function sph2d_upload_state(ctx, n, pos, vel, rho, mass, energy, h, ptype, ld) &
        bind(C, name="sph2d_upload_state") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n, ld
      real(c_double), intent(in) :: pos(ld, 2), vel(ld, 2), rho(ld), mass(ld), energy(ld), h(ld)
      integer(c_int32_t), intent(in) :: ptype(ld)
      integer(c_int) :: rc
    end function

    function sph2d_download_state(ctx, n, pos, vel, rho, press, c, mass, ptype, ld) &
        bind(C, name="sph2d_download_state") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n, ld
      real(c_double), intent(out) :: pos(ld, 2), vel(ld, 2), rho(ld), press(ld), c(ld), mass(ld)
      integer(c_int32_t), intent(out) :: ptype(ld)
      integer(c_int) :: rc
    end function

    ! compute_kernel(dim, distance, position_difference, average_smooth_length, w, temp_dwdx)
    function sph2d_compute_kernel(dim, distance, position_difference, average_smooth_length, w, &
        temp_dwdx) bind(C, name="sph2d_compute_kernel") result(rc)
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), intent(in) :: dim
      real(c_double), intent(in) :: distance, position_difference(2), average_smooth_length
      real(c_double), intent(out) :: w, temp_dwdx(2)
      integer(c_int) :: rc
    end function

    ! status_equation(particle_rho, particle_energy, particle_type, particle_press, particle_c)
    function sph2d_status_equation(particle_rho, particle_energy, particle_type, particle_press, &
        particle_c) bind(C, name="sph2d_status_equation") result(rc)
      import :: c_int, c_int32_t, c_double
      real(c_double), intent(in) :: particle_rho, particle_energy
      integer(c_int32_t), intent(in) :: particle_type
      real(c_double), intent(out) :: particle_press, particle_c
      integer(c_int) :: rc
    end function

    function sph2d_distribute_period_boundary(ctx) bind(C, name="sph2d_distribute_period_boundary") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function sph2d_devide_block(ctx, i_time_step) bind(C, name="sph2d_devide_block") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: i_time_step
      integer(c_int) :: rc
    end function

    function sph2d_single_step_compute(ctx, i_time_step) bind(C, name="sph2d_single_step_compute") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: i_time_step
      integer(c_int) :: rc
    end function

    ! single_step_compute routine by routine (single_step_compute.f90:62-171): the
    ! same call sequence as the reference's find_particle_pair (:93),
    ! Define_freesurface_and_Normailized_derivates (:108) and Solver_* (:137)
    function sph2d_find_particle_pair(ctx, i_time_step) bind(C, name="sph2d_find_particle_pair") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: i_time_step
      integer(c_int) :: rc
    end function

    ! pair_i/pair_j (1-based local indices) in find_particle_pair.f90's order
    function sph2d_download_pair_list(ctx, capacity, pair_i, pair_j, pair_number) &
        bind(C, name="sph2d_download_pair_list") result(rc)
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: capacity
      integer(c_int32_t), intent(out) :: pair_i(*), pair_j(*)
      integer(c_int64_t), intent(out) :: pair_number
      integer(c_int) :: rc
    end function

    function sph2d_define_freesurface(ctx) bind(C, name="sph2d_define_freesurface") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function sph2d_solver(ctx, variant, sps_on) bind(C, name="sph2d_solver") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: variant, sps_on
      integer(c_int) :: rc
    end function

    ! Reinitialize_particle_density, Shift_particle_position, Refresh_domain (time_iteration.f90:1064-1078)
    function sph2d_reinitialize_particle_density(ctx, i_time_step) &
        bind(C, name="sph2d_reinitialize_particle_density") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: i_time_step
      integer(c_int) :: rc
    end function

    function sph2d_shift_particle_position(ctx, i_time_step) bind(C, name="sph2d_shift_particle_position") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: i_time_step
      integer(c_int) :: rc
    end function

    function sph2d_refresh_domain(ctx, i_time_step) bind(C, name="sph2d_refresh_domain") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: i_time_step
      integer(c_int) :: rc
    end function

    ! time_iteration.f90's stage updates: scheme 2 (predictor 1, corrector 2) or 3 (K_1..K_3, final 4)
    function sph2d_stage_update(ctx, scheme, stage) bind(C, name="sph2d_stage_update") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: scheme, stage
      integer(c_int) :: rc
    end function

    ! drhodt(ld) and particle_acceleration(ld,2) in the reference's global order
    function sph2d_download_rates(ctx, n, drhodt, acc, ld) bind(C, name="sph2d_download_rates") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n, ld
      real(c_double), intent(out) :: drhodt(ld), acc(ld, 2)
      integer(c_int) :: rc
    end function

    ! Reverse of the last evaluation (no reference counterpart): weights of drho/dt,
    ! dv/dt and the XSPH velocity in, d(weighted outputs)/d(state) out, gid order.
    function sph2d_single_step_compute_vjp(ctx, n, lam_drhodt, lam_acc, lam_avg, ld, grad_pos, grad_vel, &
        grad_rho) bind(C, name="sph2d_single_step_compute_vjp") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n, ld
      real(c_double), intent(in) :: lam_drhodt(ld), lam_acc(ld, 2)
      ! absent optional dummies of a bind(C) interface pass NULL, as the header allows
      real(c_double), intent(in), optional :: lam_avg(ld, 2)
      real(c_double), intent(out), optional :: grad_pos(ld, 2), grad_vel(ld, 2), grad_rho(ld)
      integer(c_int) :: rc
    end function
    ! the same reverse on device arrays (c_ptr of device memory, (ld,2) layouts; NULL
    ! for lam_avg or an unwanted gradient)
    function sph2d_single_step_compute_vjp_device(ctx, n, lam_drhodt, lam_acc, lam_avg, ld, grad_pos, &
        grad_vel, grad_rho) bind(C, name="sph2d_single_step_compute_vjp_device") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx, lam_drhodt, lam_acc, lam_avg, grad_pos, grad_vel, grad_rho
      integer(c_int32_t), value :: n, ld
      integer(c_int) :: rc
    end function

    function sph2d_time_step(ctx, i_time_step) bind(C, name="sph2d_time_step") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: i_time_step
      integer(c_int) :: rc
    end function

    function sph2d_run(ctx, first_step, nsteps) bind(C, name="sph2d_run") result(rc)
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: first_step, nsteps
      integer(c_int) :: rc
    end function

    function sph2d_sample(ctx, out14) bind(C, name="sph2d_sample") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out14(14)
      integer(c_int) :: rc
    end function

    ! Reorder_Particle_Index (MS) on device arrays: x, y, ptype and the columns are
    ! c_ptr device addresses (e.g. from a GPU-resident driver); f64_cols/i32_cols are
    ! arrays of such addresses; counts(6) comes back on the host.
    function sph2d_reorder_particle_index(n_total, n_fluid, box_x, box_y, x, y, ptype, n_f64, f64_cols, &
                                          n_i32, i32_cols, perm, counts, stream) &
        bind(C, name="sph2d_reorder_particle_index") result(rc)
      import :: c_int, c_int32_t, c_double, c_ptr
      integer(c_int32_t), value :: n_total, n_fluid, n_f64, n_i32
      real(c_double), value :: box_x, box_y
      type(c_ptr), value :: x, y, ptype, perm, stream
      type(c_ptr), intent(in) :: f64_cols(*), i32_cols(*)
      integer(c_int32_t), intent(out) :: counts(6)
      integer(c_int) :: rc
    end function

    ! Output and restart (the reference's file formats). Paths are C strings:
    ! pass trim(name)//c_null_char.
    function sph2d_download_output_fields(ctx, n, vort_z, ana, ld, step) &
        bind(C, name="sph2d_download_output_fields") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n, ld
      real(c_double), intent(out) :: vort_z(ld), ana(ld, 4)
      integer(c_int32_t), intent(out) :: step
      integer(c_int) :: rc
    end function

    function sph2d_output_for_tecplot(ctx, i_time_step, video_path, surface_path, binary_path) &
        bind(C, name="sph2d_output_for_tecplot") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_char
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: i_time_step
      character(kind=c_char), intent(in) :: video_path(*), surface_path(*), binary_path(*)
      integer(c_int) :: rc
    end function

    function sph2d_output_for_paraview(ctx, i_time_step, path) bind(C, name="sph2d_output_for_paraview") result(rc)
      import :: c_int, c_int32_t, c_ptr, c_char
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: i_time_step
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int) :: rc
    end function

    function sph2d_write_final_results(path, n, pos, vel, rho, press, ld) &
        bind(C, name="sph2d_write_final_results") result(rc)
      import :: c_int, c_int32_t, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: n, ld
      real(c_double), intent(in) :: pos(ld, 2), vel(ld, 2), rho(ld), press(ld)
      integer(c_int) :: rc
    end function

    function sph2d_write_latest_results_zone(path, i_time_step, n, pos, vel, press, rho, mass, energy, c, &
                                             initial_type, wavemaker_layer, ld) &
        bind(C, name="sph2d_write_latest_results_zone") result(rc)
      import :: c_int, c_int32_t, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: i_time_step, n, ld
      real(c_double), intent(in) :: pos(ld, 2), vel(ld, 2), press(ld), rho(ld), mass(ld), energy(ld), c(ld)
      integer(c_int32_t), intent(in) :: initial_type(ld), wavemaker_layer(ld)
      integer(c_int) :: rc
    end function

    function sph2d_read_latest_results(path, zone, capacity, n_out, pos, vel, press, rho, mass, energy, c, &
                                       initial_type, wavemaker_layer, step_out, ld) &
        bind(C, name="sph2d_read_latest_results") result(rc)
      import :: c_int, c_int32_t, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: zone, capacity, ld
      integer(c_int32_t), intent(out) :: n_out, step_out
      real(c_double), intent(out) :: pos(ld, 2), vel(ld, 2), press(ld), rho(ld), mass(ld), energy(ld), c(ld)
      integer(c_int32_t), intent(out) :: initial_type(ld), wavemaker_layer(ld)
      integer(c_int) :: rc
    end function

    function sph2d_write_sampling(dir, nondim_time, out14) bind(C, name="sph2d_write_sampling") result(rc)
      import :: c_int, c_double, c_char
      character(kind=c_char), intent(in) :: dir(*)
      real(c_double), value :: nondim_time
      real(c_double), intent(in) :: out14(14)
      integer(c_int) :: rc
    end function

    ! Sampling_subroutine.f90:21-177 at n points (pos(ld,2)), after a sampling step
    function sph2d_sampling(ctx, i_time_step, n, pos, ld, gravity, average_press, mls_press, vel) &
        bind(C, name="sph2d_sampling") result(rc)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: i_time_step, n, ld
      real(c_double), intent(in) :: pos(ld, 2)
      real(c_double), value :: gravity
      real(c_double), intent(out) :: average_press(n), mls_press(n), vel(ld, 2)
      integer(c_int) :: rc
    end function

    ! Sampling_subroutine.f90:246-279: wave front and wave gauge elevations
    function sph2d_wave_probes(ctx, n_probes, probe_x, wave_front, elevation) &
        bind(C, name="sph2d_wave_probes") result(rc)
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n_probes
      real(c_double), intent(in) :: probe_x(*)
      real(c_double), intent(out) :: wave_front, elevation(*)
      integer(c_int) :: rc
    end function

    ! Sampling_subroutine.f90:300-322
    function sph2d_kalman_filter(n, current_time_number, measured, last_x, last_p, filtered) &
        bind(C, name="sph2d_kalman_filter") result(rc)
      import :: c_int, c_int32_t, c_double
      integer(c_int32_t), value :: n, current_time_number
      real(c_double), intent(in) :: measured(n)
      real(c_double), intent(inout) :: last_x(n), last_p(n)
      real(c_double), intent(out) :: filtered(n)
      integer(c_int) :: rc
    end function

    ! write(unit,*) of n REAL(8) values, appended (Wave_front_results.dat, wave gauges)
    function sph2d_write_list_directed(path, n, values) bind(C, name="sph2d_write_list_directed") result(rc)
      import :: c_int, c_int32_t, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: n
      real(c_double), intent(in) :: values(n)
      integer(c_int) :: rc
    end function

    ! write(unit,'(20F20.8)') of n values, appended (pressure probe files)
    function sph2d_write_f20_8(path, n, values) bind(C, name="sph2d_write_f20_8") result(rc)
      import :: c_int, c_int32_t, c_double, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: n
      real(c_double), intent(in) :: values(n)
      integer(c_int) :: rc
    end function

    ! Wavemaker_Subroutine.f90:66-138 and wavemaker_plate_motion.f90: the paddle at step i_time_step
    function sph2d_wavemaker_motion(w, i_time_step, position, velocity, acceleration) &
        bind(C, name="sph2d_wavemaker_motion") result(rc)
      import :: sph2d_wavemaker, c_int, c_int32_t, c_double
      type(sph2d_wavemaker), intent(in) :: w
      integer(c_int32_t), value :: i_time_step
      real(c_double), intent(out) :: position, velocity, acceleration
      integer(c_int) :: rc
    end function

    ! Wavemaker_Subroutine.f90:140-215 (paddle position and velocity from sph2d_wavemaker_motion)
    function sph2d_wavemaker_update(ctx, position, velocity) bind(C, name="sph2d_wavemaker_update") result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: position, velocity
      integer(c_int) :: rc
    end function

    function sph2d_last_error() bind(C, name="sph2d_last_error") result(msg)
      import :: c_ptr
      type(c_ptr) :: msg
    end function
  end interface
end module sph2d_interface
