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
!
! Device-authoritative mode (environment GPE_RESIDENT=1, not with
! diagnostic): the field stays in HBM across calls.  The unchanged
! driver copies psi%new into psi%old and calls solver() every step
! (gpe.f90:167-176, 230), but it reads psi only for its periodic I/O
! (gpe.f90:178-228, 234, 245-247) and at the end of the run.  So the module
! uploads var_in once, skips the upload while var_in is the field it
! returned, and writes var_out only when the driver will read psi before the
! next call.  It predicts that from the driver's own schedule: the
! save_rate test, the save_rate2/save_rate3/p_save thresholds with mirrors
! of the driver's counters n, m, ps (gpe.f90:127-129 and their
! increments), the end_time test of the loop, and the SAVE and RUNNING
! files.  Two reads it cannot predict: a RUNNING file deleted after this
! call returns (then end_state writes psi%new, io.f90:831) and a SAVE file
! created after it (save_run, io.f90:1558).  A call of solver_sync(in_var)
! at the top of end_state and inside save_run's `if (save3d == 1)` makes
! those exact.  get_norm(psi%new) runs every step (gpe.f90:244); with
! `if (solver_norm(in_var, norm)) return` at its top it takes the norm of
! the device field instead of the stale host copy.
!
! Imaginary time (eqn 4) changes psi between calls: get_norm decides the
! switch to real time and renormalise rescales psi (gpe.f90:243-273).  The
! field stays resident there only once get_norm's hook has been called and,
! when renorm is set, renormalise's too (`if (solver_renormalise(var,
! norm)) return` at its top), and not with imprint_vl.  solver_norm then
! repeats the driver's test abs((norm-prev_norm)/prev_norm) < 1e-12 with the
! norms it returned and writes psi%new back in the iteration that switches
! (idl_surface may read it there), or whenever the previous norm is not
! its own.
module solve
  use parameters
  use iso_c_binding
  implicit none

  private
  public :: solver, solver_norm, solver_sync, solver_renormalise, solver_transfer_bytes

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
    function gpe_solver_step_ex(ctx, vin, halo, vout, st, flags) bind(C, name='gpe_solver_step_ex') result(rc)
      import :: c_ptr, c_int, c_int32_t, gpe_step_state_t
      type(c_ptr), value :: ctx, vin, vout
      integer(c_int32_t), value :: halo, flags
      type(gpe_step_state_t), intent(inout) :: st
      integer(c_int) :: rc
    end function gpe_solver_step_ex
    function gpe_sim_download(ctx, psi, host_type) bind(C, name='gpe_sim_download') result(rc)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx, psi
      integer(c_int32_t), value :: host_type
      integer(c_int) :: rc
    end function gpe_sim_download
    function gpe_sim_norm(ctx, norm) bind(C, name='gpe_sim_norm') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: norm
      integer(c_int) :: rc
    end function gpe_sim_norm
    function gpe_transfer_bytes(ctx, h2d, d2h) bind(C, name='gpe_transfer_bytes') result(rc)
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: h2d, d2h
      integer(c_int) :: rc
    end function gpe_transfer_bytes
    function gpe_sim_renormalise(ctx, nn, norm) bind(C, name='gpe_sim_renormalise') result(rc)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: nn, norm
      integer(c_int) :: rc
    end function gpe_sim_renormalise
    function gpe_tune_load(ctx, path, applied) bind(C, name='gpe_tune_load') result(rc)
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx, path
      integer(c_int32_t), intent(out) :: applied
      integer(c_int) :: rc
    end function gpe_tune_load
    subroutine gpe_ctx_destroy(ctx) bind(C, name='gpe_ctx_destroy')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine gpe_ctx_destroy
    function gpe_last_error() bind(C, name='gpe_last_error') result(msg)
      import :: c_ptr
      type(c_ptr) :: msg
    end function gpe_last_error
  end interface

  integer(c_int32_t), parameter :: step_in_resident = 1, step_out_deferred = 2  ! include/gpe_hip.h

  type(c_ptr), save :: ctx = c_null_ptr
  logical, save :: ctx_real_time = .true.  ! real_time the context was made for
  ! device-authoritative mode
  logical, save :: resident = .false.   ! GPE_RESIDENT=1 at the first call
  logical, save :: dev_holds = .false.  ! the context holds the current field
  logical, save :: host_stale = .false. ! held_out is older than the device field
  type(c_ptr), save :: held_out = c_null_ptr
  integer, save :: n_io = 0, m_io = 0, ps_io = 0  ! mirrors of gpe.f90's n, m, ps
  ! imaginary time: resident once get_norm's hook (and renormalise's, when
  ! renorm is set) have been called, so both act on the device field
  logical, save :: norm_hook = .false., renorm_hook = .false.
  real (pr), save :: last_norm = 0.0_pr  ! the norm solver_norm last returned
  logical, save :: have_last = .false.   ! ... in the iteration before this one
  integer(c_int64_t), save :: gone_h2d = 0, gone_d2h = 0  ! bytes moved by destroyed contexts

  contains

  subroutine solver(var_in, var_out)
    ! One time step of the configured scheme (solve.f90:25-53)
    use error
    use parameters
    implicit none

    complex (pr), dimension(0:nx1,js-2:je+2,ks-2:ke+2), intent(in), target :: var_in
    ! inout: in device-authoritative mode var_out keeps its contents on the
    ! steps after which the driver does not read it
    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(inout), target :: var_out
    type(gpe_step_state_t) :: st
    integer(c_int) :: rc
    integer(c_int32_t) :: flags
    integer(c_int64_t) :: h2d_now, d2h_now
    logical :: use_res

    ! gpe.f90 switches from imaginary to real time once the norm has
    ! converged (gpe.f90:250-264); the rk45 control of a context follows the
    ! real_time it was created with, so a switch makes a new one
    if (c_associated(ctx) .and. (real_time .neqv. ctx_real_time)) then
      ! solver_norm writes the field back in the iteration that switches
      if (dev_holds .and. host_stale) call emergency_stop('ERROR: the device-held field was not '// &
        'written back before the switch to real time (get_norm must call solver_norm).')
      call solver_transfer_bytes(h2d_now, d2h_now)
      gone_h2d = h2d_now
      gone_d2h = d2h_now
      call gpe_ctx_destroy(ctx)
      ctx = c_null_ptr
      dev_holds = .false.
      host_stale = .false.
    end if
    if (.not. c_associated(ctx)) call create_context()

    ! imaginary time changes psi between steps through get_norm and
    ! renormalise (gpe.f90:243-273): resident only with both hooked, and
    ! not with imprint_vl
    use_res = resident .and. .not. diagnostic .and. (real_time .or. (eqn_to_solve == 4 .and. norm_hook .and. &
      (renorm_hook .or. .not. renorm) .and. .not. imprint_vl))
    flags = 0
    if (use_res) then
      if (.not. dev_holds) then
        ! first resident step (or the first after imaginary time): the
        ! driver's counters as gpe.f90:127-129 (and 258-260) set them
        ps_io = int(t/p_save)
        n_io = int(t/save_rate2)
        m_io = int(t/save_rate3)
      end if
      ! the increments the driver made in this iteration (gpe.f90:188-228)
      if (t+im_t >= save_rate3*m_io) m_io = m_io+1
      if (t+im_t >= save_rate2*n_io) n_io = n_io+1
      if (t+im_t >= p_save*ps_io) ps_io = ps_io+1
      flags = step_out_deferred
      if (dev_holds) flags = flags + step_in_resident
    else
      dev_holds = .false.
      host_stale = .false.
    end if

    st%dt_re = real(dt, c_double)
    st%dt_im = real(aimag(dt), c_double)
    st%t = real(t, c_double)
    st%im_t = real(im_t, c_double)
    st%p = int(p, c_int32_t)
    st%errmax = 0.0_c_double
    st%n_rejects = 0

    rc = gpe_solver_step_ex(ctx, c_loc(var_in), 2_c_int32_t, c_loc(var_out), st, flags)
    if (rc /= 0) call emergency_stop('ERROR: '//last_error())

    dt = cmplx(st%dt_re, st%dt_im, pr)
    t = real(st%t, pr)
    im_t = real(st%im_t, pr)

    if (use_res) then
      dev_holds = .true.
      held_out = c_loc(var_out)
      host_stale = .true.
      if (host_reads_next()) call solver_sync(var_out)
    end if

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

  logical function host_reads_next()
    ! Will the driver read psi before its next solver() call?  The tests of
    ! gpe.f90's loop, evaluated with the time, dt and counters it will use.
    use parameters
    implicit none

    logical :: file_exists

    host_reads_next = .true.
    ! the loop ends (gpe.f90:136): end_state, idl_surface, save_surface
    if (t+im_t > end_time) return
    ! save_norm now (245-247); save_energy, save_time, save_linelength at
    ! the start of the next iteration (178-186)
    if (modulo(t+im_t, abs(dt)*save_rate) < abs(dt)) return
    ! condensed_particles, pdf, vcf (188-197)
    if (t+im_t >= save_rate3*m_io) return
    ! zeros, surfaces, averages (199-221) and diag (236-240)
    if (t+im_t >= save_rate2*n_io) return
    ! periodic end_state (224-228)
    if (t+im_t >= p_save*ps_io) return
    ! save_run (234, io.f90:1558-1587)
    inquire (file='SAVE', exist=file_exists)
    if (file_exists) return
    ! a stop request (139-159): end_state(psi%new, 1)
    inquire (file='RUNNING', exist=file_exists)
    if (.not. file_exists) return
    host_reads_next = .false.

    return
  end function host_reads_next

! ***************************************************************************

  subroutine solver_sync(var)
    ! Bring var up to date if it is the field a device-authoritative step
    ! left stale (a no-op otherwise).  For a host reader the module cannot
    ! predict: call it at the top of end_state (io.f90:831) and inside
    ! save_run's `if (save3d == 1)` (io.f90:1578).
    use error
    use parameters
    implicit none

    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(inout), target :: var
    integer(c_int) :: rc
    integer(c_int32_t) :: ht

    if (.not. host_stale) return
    if (.not. c_associated(held_out, c_loc(var))) return
    ht = 0
    if (kind(var) /= kind((1.0d0, 1.0d0))) ht = 1
    rc = gpe_sim_download(ctx, c_loc(var), ht)
    if (rc /= 0) call emergency_stop('ERROR: '//last_error())
    host_stale = .false.

    return
  end subroutine solver_sync

! ***************************************************************************

  logical function solver_norm(in_var, norm)
    ! get_norm (variables.f90:554-576) of the device field when in_var is
    ! the field a device-authoritative step returned: .true. and norm set.
    ! .false. otherwise (get_norm then integrates in_var on the host).
    use error
    use parameters
    implicit none

    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(in), target :: in_var
    real (pr), intent(out) :: norm
    real(c_double) :: v
    integer(c_int) :: rc

    norm_hook = .true.
    solver_norm = .false.
    norm = 0.0_pr
    if (.not. dev_holds .or. .not. c_associated(held_out, c_loc(in_var))) then
      have_last = .false.
      return
    end if
    rc = gpe_sim_norm(ctx, v)
    if (rc /= 0) call emergency_stop('ERROR: '//last_error())
    norm = real(v, pr)
    solver_norm = .true.
    if (.not. real_time) then
      ! gpe.f90:245-264 switches to real time (or stops) once
      ! abs((norm-prev_norm)/prev_norm) < 1e-12, and then may read psi%new:
      ! write it back in that iteration (and whenever prev_norm is unknown)
      if (.not. have_last) then
        call write_back(c_loc(in_var))
      else if (abs((norm-last_norm)/last_norm) < 1e-12_pr) then
        call write_back(c_loc(in_var))
      end if
    end if
    last_norm = norm
    have_last = .true.

    return
  end function solver_norm

! ***************************************************************************

  logical function solver_renormalise(var, norm)
    ! renormalise (variables.f90:715-726) of the device field when var is
    ! the field a device-authoritative step returned: .true. and the field
    ! (and var, if it was current) renormalised.  .false. otherwise
    ! (renormalise then works on var on the host).  Call it at the top of
    ! renormalise:  if (solver_renormalise(var, norm)) return
    use error
    use parameters
    implicit none

    complex (pr), dimension(0:nx1,js:je,ks:ke), intent(inout), target :: var
    real (pr), intent(in) :: norm
    integer(c_int) :: rc
    logical :: was_current

    renorm_hook = .true.
    solver_renormalise = .false.
    if (.not. dev_holds) return
    if (.not. c_associated(held_out, c_loc(var))) return
    rc = gpe_sim_renormalise(ctx, real(nn, c_double), real(norm, c_double))
    if (rc /= 0) call emergency_stop('ERROR: '//last_error())
    was_current = .not. host_stale
    host_stale = .true.
    if (was_current) call solver_sync(var)
    solver_renormalise = .true.

    return
  end function solver_renormalise

! ***************************************************************************

  subroutine write_back(addr)
    ! Download the device field into the host array at addr if that is the
    ! stale field (solver_sync for a caller that holds only the address)
    use error
    use parameters
    implicit none

    type(c_ptr), intent(in) :: addr
    integer(c_int) :: rc
    integer(c_int32_t) :: ht

    if (.not. host_stale) return
    if (.not. c_associated(held_out, addr)) return
    ht = 0
    if (pr /= kind(1.0d0)) ht = 1
    rc = gpe_sim_download(ctx, addr, ht)
    if (rc /= 0) call emergency_stop('ERROR: '//last_error())
    host_stale = .false.

    return
  end subroutine write_back

! ***************************************************************************

  subroutine solver_transfer_bytes(to_device, to_host)
    ! Field bytes moved host -> device and device -> host so far
    implicit none

    integer(c_int64_t), intent(out) :: to_device, to_host
    integer(c_int) :: rc

    to_device = 0
    to_host = 0
    if (c_associated(ctx)) rc = gpe_transfer_bytes(ctx, to_device, to_host)
    ! and what the contexts replaced at a switch to real time moved
    to_device = to_device + gone_h2d
    to_host = to_host + gone_d2h

    return
  end subroutine solver_transfer_bytes

! ***************************************************************************

  subroutine create_context()
    ! Pack the run parameters into the C struct and create the device context
    use error
    use parameters
    implicit none

    type(gpe_params_t) :: prm
    integer(c_int) :: rc
    integer(c_int32_t) :: applied
    character(len=8) :: flag
    integer :: st

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
    ! stored launch configuration of this problem ($GPE_TUNING_FILE, else
    ! tuning.json beside the library), if any: speed only
    rc = gpe_tune_load(ctx, c_null_ptr, applied)
    if (rc /= 0) call emergency_stop('ERROR: '//last_error())
    ctx_real_time = real_time

    call get_environment_variable('GPE_RESIDENT', flag, status=st)
    resident = st == 0 .and. trim(flag) == '1'

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
