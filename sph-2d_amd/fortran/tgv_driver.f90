! tgv_driver.f90 -- a Fortran caller of the GPU step through sph2d_interface:
! the shape of time_iteration.f90's loop with the reference's argument lists.
! Usage: tgv_driver [n_side] [nsteps]; prints one line per 100 steps
! (step, KE/KE0, L2 of the u error) and a final "ok".
program tgv_driver
  use iso_c_binding
  use sph2d_interface
  implicit none
  type(sph2d_params) :: prm
  type(c_ptr) :: ctx
  integer(c_int) :: rc
  integer :: n_side, nsteps, it, nargs
  character(len=32) :: arg
  real(c_double) :: out(14), w, dwdx(2), pd(2), r, h, rho, e, p, c
  integer(c_int32_t) :: dim, ptype

  n_side = 100
  nsteps = 200
  nargs = command_argument_count()
  if (nargs >= 1) then
    call get_command_argument(1, arg)
    read(arg, *) n_side
  end if
  if (nargs >= 2) then
    call get_command_argument(2, arg)
    read(arg, *) nsteps
  end if

  ! scalar entry points, called with the reference's argument lists
  dim = 2
  h = 0.012d0
  pd = (/ 0.006d0, 0.008d0 /)
  r = sqrt(pd(1)*pd(1) + pd(2)*pd(2))
  rc = sph2d_compute_kernel(dim, r, pd, h, w, dwdx)
  call check(rc)
  rho = 1001.0d0
  e = 0.0d0
  ptype = 2
  rc = sph2d_status_equation(rho, e, ptype, p, c)
  call check(rc)
  write(*, '(a, 3es24.16)') 'compute_kernel ', w, dwdx
  write(*, '(a, 2es24.16)') 'status_equation ', p, c

  rc = sph2d_default_params(prm)
  call check(rc)
  prm%nx = n_side
  prm%ny = n_side
  prm%dx = 1.0d0 / n_side
  prm%h = 1.20d0 * prm%dx
  prm%max_shift = 0.25d0 * prm%dx
  prm%dt = 1.0d-4 * (100.0d0 / n_side)
  prm%integrator = SPH2D_INTEGRATOR_RK
  rc = sph2d_ctx_create(prm, n_side * n_side, ctx)
  call check(rc)
  rc = sph2d_init_tgv(ctx)
  call check(rc)
  do it = 1, nsteps
    rc = sph2d_time_step(ctx, it)
    call check(rc)
    if (mod(it, 100) == 0) then
      rc = sph2d_sample(ctx, out)
      call check(rc)
      write(*, '(i8, 2es24.16)') it, out(1), out(5)
    end if
  end do
  rc = sph2d_ctx_destroy(ctx)
  write(*, '(a)') 'ok'

contains

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
end program tgv_driver
