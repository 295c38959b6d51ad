! solve.f90 -- drop-in module solve for the reference's Fortran driver.
!
! Same public interface as the reference's module solve (solve.f90:15-53):
! `call solver(var_in, var_out)`, called once per time step by gpe.f90:230.
! The step is evaluated by the MI355X library libgpehip through its C ABI
! (include/gpe_hip.h): gpe_solver_step replaces euler/rk2/rk4/rkqs
! (solve.f90:57-378) and get_rhs (382-440).  The module reads and writes
! the same globals of module parameters as the reference's solver: scheme,
! eqn_to_solve, bcs, order, real_time, g, mu, omx/omy/omz, Urhs, diss_amp,
! the grid (xl.., dx..), eps/safety/dt_increase/dt_decrease in; t, im_t,
! dt and errcon out; and rk45 appends timestep.dat as rkqs does
! (solve.f90:234-241).  Errors end the run through emergency_stop
! (error.f90:24-46), as the reference does for an unknown scheme.
!
! One rank only (nyprocs = nzprocs = 1): boundaries are applied in the GPU
! kernels, so the ghost planes of var_in are not read.
module solve
  use parameters
  use iso_c_binding
  implicit none

  private
  public :: solver

  ! mirrors struct gpe_params (include/gpe_hip.h)
  type, bind(C) :: gpe_params_t
    integer(c_int32_t) :: nx, ny, nz, scheme, eqn, bcs, order, precision, real_time, reserved
    real(c_double) :: xl, yl, zl, dx, dy, dz, g, mu, omx, omy, omz, urhs, diss_amp
    real(c_double) :: eps, safety, dt_increase, dt_decrease
  end type gpe_params_t

  ! mirrors struct gpe_step_state
  type, bind(C) :: gpe_step_state_t
    real(c_double) :: dt_re, dt_im, t, im_t, errmax
    integer(c_int32_t) :: n_rejects, p
  end type gpe_step_state_t

  interface
    function gpe_ctx_create(prm, ctx) bind(C, name='gpe_ctx_create') result(rc)
      import :: gpe_params_t, c_ptr, c_int
      type(gpe_params_t), intent(in) :: prm
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: rc
    end function gpe_ctx_create
    function gpe_solver_step(ctx, vin, halo, vout, st) bind(C, name='gpe_solver_step') result(rc)
      import :: c_ptr, c_int, c_int32_t, gpe_step_state_t
      type(c_ptr), value :: ctx, vin, vout
      integer(c_int32_t), value :: halo
      type(gpe_step_state_t), intent(inout) :: st
      integer(c_int) :: rc
    end function gpe_solver_step
    function gpe_last_error() bind(C, name='gpe_last_error') result(msg)
      import :: c_ptr
      type(c_ptr) :: msg
    end function gpe_last_error
  end interface

  type(c_ptr), save :: ctx = c_null_ptr

  contains

  subroutine solver(var_in, var_out)
    ! One time step of the configured scheme (solve.f90:25-53)
    use error
    use parameters
    implicit none

    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in), target :: var_in
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(out), target :: var_out
    type(gpe_step_state_t) :: st
    integer(c_int) :: rc

    if (.not. c_associated(ctx)) call create_context()

    st%dt_re = real(dt, c_double)
    st%dt_im = real(aimag(dt), c_double)
    st%t = real(t, c_double)
    st%im_t = real(im_t, c_double)
    st%p = int(p, c_int32_t)
    st%errmax = 0.0_c_double
    st%n_rejects = 0

    rc = gpe_solver_step(ctx, c_loc(var_in), 2_c_int32_t, c_loc(var_out), st)
    if (rc /= 0) call emergency_stop('ERROR: '//last_error())

    dt = cmplx(st%dt_re, st%dt_im, pr)
    t = real(st%t, pr)
    im_t = real(st%im_t, pr)

    if (scheme == 'rk45') then
      errcon = (5.0_pr/safety)**(1.0_pr/dt_increase)
      if (myrank == 0) then
        if (mod(p, save_rate) == 0) then
          open (11, status='unknown', position='append', file='timestep.dat')
          write (11, '(3e17.9,i10)') t, im_t, abs(dt), p
          close (11)
        end if
      end if
    end if

    return
  end subroutine solver

! ***************************************************************************

  subroutine create_context()
    ! Pack the run parameters into the C struct and create the device context
    use error
    use parameters
    implicit none

    type(gpe_params_t) :: prm
    integer(c_int) :: rc

    if (nyprocs*nzprocs /= 1) then
      call emergency_stop('ERROR: the GPU solver runs on one rank (nyprocs = nzprocs = 1).')
    end if

    prm%nx = nx
    prm%ny = ny
    prm%nz = nz
    select case (scheme)
      case ('euler')
        prm%scheme = 1
      case ('rk2')
        prm%scheme = 2
      case ('rk4')
        prm%scheme = 4
      case ('rk45')
        prm%scheme = 45
      case default
        call emergency_stop('ERROR: Unrecognised time stepping scheme.')
    end select
    prm%eqn = eqn_to_solve
    prm%bcs = bcs
    prm%order = order
    if (kind(dx) == kind(1.0d0)) then
      prm%precision = 0
    else
      prm%precision = 1
    end if
    prm%real_time = merge(1, 0, real_time)
    prm%reserved = 0
    prm%xl = xl
    prm%yl = yl
    prm%zl = zl
    prm%dx = dx
    prm%dy = dy
    prm%dz = dz
    prm%g = g
    prm%mu = mu
    prm%omx = omx
    prm%omy = omy
    prm%omz = omz
    prm%urhs = Urhs
    prm%diss_amp = diss_amp
    prm%eps = eps
    prm%safety = safety
    prm%dt_increase = dt_increase
    prm%dt_decrease = dt_decrease

    rc = gpe_ctx_create(prm, ctx)
    if (rc /= 0) call emergency_stop('ERROR: '//last_error())

    return
  end subroutine create_context

! ***************************************************************************

  function last_error()
    ! Message of the last failed library call
    implicit none

    character(len=:), allocatable :: last_error
    character(kind=c_char), pointer :: chars(:)
    type(c_ptr) :: msg
    integer :: n

    msg = gpe_last_error()
    last_error = ''
    if (.not. c_associated(msg)) return
    call c_f_pointer(msg, chars, [512])
    n = 0
    do while (n < 512)
      if (chars(n+1) == c_null_char) exit
      n = n+1
    end do
    allocate(character(len=n) :: last_error)
    last_error = transfer(chars(1:n), last_error)

    return
  end function last_error

end module solve
