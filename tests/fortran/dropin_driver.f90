! dropin_driver.f90 -- drives the drop-in module solve the way gpe.f90 does.
!
! TEST HARNESS.  Built by oracle/Makefile (target dropin) against the
! reference's unchanged parameters.f90 and error.f90 and the drop-in
! gpe_amd/fortran/solve.f90.  It reads the reference's own namelists
! run_params, eqn_params and (if present) io_params (parameters.f90:40-76)
! from a run.in-style file, forms the grid as get_grid does (ic.f90:43-66),
! reads the initial state in the restart format of end_state
! (io.f90:843-853), and runs the time loop of gpe.f90:136-280 while
! t+im_t <= end_time: the RUNNING test, the copy of psi%new into psi%old,
! the I/O tests in the driver's order, solver(), the SAVE test of save_run,
! get_norm, p = p+1.  Each place where gpe.f90 would read psi for I/O
! appends a record (p, what, t, the field it would read) to io.bin instead,
! so two runs can be compared record by record.  get_norm is taken with
! the drop-in's solver_norm hook where that applies (the norms go to
! norms.txt), and every end_state record first calls solver_sync, as the
! hooked end_state would.  The final state is written in the end_state
! format.
!
! Environment (test hooks): DRIVER_STOP_AT=p deletes RUNNING at the top of
! iteration p (a user stopping the run); DRIVER_SAVE_AT=p creates SAVE just
! before solver() in iteration p; DRIVER_REAL_AT=p switches an
! imaginary-time run to real time at the top of iteration p the way
! gpe.f90:250-264 does once the norm has converged (real_time = .true.,
! im_t = 0, dt = (|Im dt|, |Re dt|)).  DRIVER_IMAG=1 runs the imaginary-
! time part of the loop instead (eqn 4): get_norm before the loop and
! renormalise if renorm is set (gpe.f90:112-113); each iteration the
! relative norm change, the switch to real time once it is below 1e-12
! (recording psi%new as idl_surface would, record 10) and renormalise
! (gpe.f90:243-276), with get_norm and renormalise hooked as a maintainer
! hooks variables.f90 (solver_norm, solver_renormalise; the host forms,
! restated below, when the hooks decline).
!
! Usage: dropin_driver <run.in> <state_in> <state_out>
program dropin_driver
  use parameters
  use solve
  use iso_c_binding, only : c_int64_t
  implicit none

  complex (pr), allocatable :: old(:,:,:), new(:,:,:)
  character(len=512) :: fin_nml, fin, fout
  integer :: u, io, nx_in, ny_in, nz_in, p_in, n, m, ps, stop_at, save_at, real_at, ios, imag, uim
  real (pr) :: t_in, im_t_in, norm, prev_norm, relnorm
  complex (pr) :: dt_in
  logical :: run_exist, save_exist, hooked
  integer(c_int64_t) :: h2d, d2h

  call MPI_INIT(ierr)
  call get_command_argument(1, fin_nml)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)

  ! defaults for what this harness does not read
  save_rate = 1
  save_rate2 = huge(1.0_pr)
  save_rate3 = huge(1.0_pr)
  p_save = huge(1.0_pr)
  diagnostic = .false.
  open (newunit=u, file=trim(fin_nml), status='old')
  read (u, nml=run_params)
  read (u, nml=eqn_params)
  read (u, nml=io_params, iostat=ios)
  close (u)
  stop_at = env_int('DRIVER_STOP_AT')
  save_at = env_int('DRIVER_SAVE_AT')
  real_at = env_int('DRIVER_REAL_AT')
  imag = env_int('DRIVER_IMAG')

  myrank = 0
  js = 0; je = ny-1; ks = 0; ke = nz-1
  xl = -xr; yl = -yr; zl = -zr
  dx = (xr-xl)/nx1; dy = (yr-yl)/ny1; dz = (zr-zl)/nz1
  dx2 = dx**2; dy2 = dy**2; dz2 = dz**2

  allocate(old(0:nx1,js-2:je+2,ks-2:ke+2), new(0:nx1,js:je,ks:ke))
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u) nx_in, ny_in, nz_in, p_in
  read (u) t_in, im_t_in
  read (u) dt_in
  read (u) new
  close (u)
  if (nx_in /= nx .or. ny_in /= ny .or. nz_in /= nz) stop 'grid mismatch'
  p = p_in
  t = t_in
  im_t = im_t_in
  dt = dt_in

  open (newunit=io, file='io.bin', access='stream', form='unformatted', status='replace')
  open (newunit=u, file='norms.txt', status='replace')
  open (newunit=uim, file='imag.txt', status='replace')  ! p, real_time, norm, relnorm (DRIVER_IMAG)

  ps = int(t/p_save)
  n = int(t/save_rate2)
  m = int(t/save_rate3)
  old = 0.0_pr
  prev_norm = 1.0_pr
  relnorm = 0.0_pr
  if (imag == 1) then
    call get_norm(new, prev_norm)
    if (renorm) call renormalise(new, prev_norm)
  end if
  do while (t+im_t <= end_time)
    if (p == real_at .and. .not. real_time) then
      call solver_sync(new)   ! a switch forced outside the norm test
      real_time = .true.
      im_t = 0.0_pr
      dt = cmplx(abs(aimag(dt)), abs(real(dt, pr)), pr)
    end if
    if (p == stop_at) then
      open (99, file='RUNNING')
      close (99, status='delete')
    end if
    inquire (file='RUNNING', exist=run_exist)
    if (.not. run_exist) then
      call solver_sync(new)
      call record(1, new)          ! end_state(psi%new, 1)
      exit
    end if
    old(:,js:je,ks:ke) = new
    if (modulo(t+im_t, abs(dt)*save_rate) < abs(dt)) then
      call record(2, old(:,js:je,ks:ke))   ! save_energy(t, psi%old), save_linelength
      call record(3, new)                  ! save_time(t, psi%new)
    end if
    if (t+im_t >= save_rate3*m) then
      call record(4, new)          ! condensed_particles(t, psi%new)
      m = m+1
    end if
    if (t+im_t >= save_rate2*n) then
      call record(5, new)          ! save_surface / idl_surface / average
      n = n+1
    end if
    if (t+im_t >= p_save*ps) then
      call solver_sync(new)
      call record(6, new)          ! end_state(psi%new, 0)
      ps = ps+1
    end if
    if (p == save_at) then
      open (98, file='SAVE')
      close (98)
    end if
    call solver(old, new)
    inquire (file='SAVE', exist=save_exist)
    if (save_exist) then
      call solver_sync(new)
      call record(7, new)          ! save_run: idl_surface(psi%new)
      open (98, file='SAVE')
      close (98, status='delete')
    end if
    if (imag == 1) then
      call get_norm(new, norm)
      relnorm = abs((norm-prev_norm)/prev_norm)
      write (uim, '(i10,l2,2es26.17)') p, real_time, norm, relnorm
    else
      hooked = solver_norm(new, norm)
      if (hooked) write (u, '(i10,es26.17)') p, norm
    end if
    if (modulo(t+im_t, abs(dt)*save_rate) < abs(dt)) then
      call record(8, new)          ! get_norm(psi%new) for save_norm
    end if
    if (imag == 1 .and. .not. real_time) then
      if (relnorm < 1e-12_pr) then
        real_time = .true.
        im_t = 0.0_pr
        ps = int(t/p_save)
        n = int(t/save_rate2)
        m = int(t/save_rate3)
        call record(10, new)       ! idl_surface(psi%new) with save_3d
        dt = cmplx(abs(aimag(dt)), abs(real(dt, pr)), pr)
      end if
      if (renorm) call renormalise(new, norm)
    end if
    prev_norm = norm
    p = p+1
  end do
  close (u)
  close (uim)
  if (run_exist) then
    call solver_sync(new)
    call record(9, new)            ! end_state(psi%new, 1) after the loop
  end if
  close (io)

  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) nx, ny, nz, p
  write (u) t, im_t
  write (u) dt
  write (u) new
  close (u)
  call solver_transfer_bytes(h2d, d2h)
  print '(a,i0,a,i0)', 'transfer_bytes ', h2d, ' ', d2h
  call MPI_FINALIZE(ierr)

contains

  subroutine record(what, f)
    integer, intent(in) :: what
    complex (pr), intent(in) :: f(0:nx1,js:je,ks:ke)
    write (io) p, what, t
    write (io) f
  end subroutine record

  subroutine get_norm(in_var, norm)
    ! variables.f90:554-576 with its hook: the device norm when the drop-in
    ! holds in_var, else the closed-rule integral on the host (weights
    ! 3/8, 7/6, 23/24 at each end, variables.f90:49-51, 730-840)
    complex (pr), intent(in) :: in_var(0:nx1,js:je,ks:ke)
    real (pr), intent(out) :: norm
    real (pr) :: w(0:max(nx1,ny1,nz1), 3), s, sy
    integer :: i, j, k, a, nn_a(3)

    if (solver_norm(in_var, norm)) then
      write (u, '(i10,es26.17)') p, norm
      return
    end if
    nn_a = (/ nx, ny, nz /)
    w = 1.0_pr
    do a = 1, 3
      w(0, a) = 3.0_pr/8.0_pr
      w(1, a) = 7.0_pr/6.0_pr
      w(2, a) = 23.0_pr/24.0_pr
      w(nn_a(a)-3, a) = 23.0_pr/24.0_pr
      w(nn_a(a)-2, a) = 7.0_pr/6.0_pr
      w(nn_a(a)-1, a) = 3.0_pr/8.0_pr
    end do
    norm = 0.0_pr
    do k = ks, ke
      sy = 0.0_pr
      do j = js, je
        s = 0.0_pr
        do i = 0, nx1
          s = s + w(i, 1)*abs(in_var(i,j,k))**2
        end do
        sy = sy + w(j, 2)*s*dx
      end do
      norm = norm + w(k, 3)*sy*dy
    end do
    norm = norm*dz
  end subroutine get_norm

  subroutine renormalise(var, norm)
    ! variables.f90:715-726 with its hook
    complex (pr), intent(inout) :: var(0:nx1,js:je,ks:ke)
    real (pr), intent(in) :: norm

    if (solver_renormalise(var, norm)) return
    var = sqrt(nn) * var / sqrt(norm)
  end subroutine renormalise

  integer function env_int(name)
    character(len=*), intent(in) :: name
    character(len=32) :: v
    integer :: st, e
    env_int = -1
    call get_environment_variable(name, v, status=st)
    if (st /= 0) return
    read (v, *, iostat=e) env_int
    if (e /= 0) env_int = -1
  end function env_int

end program dropin_driver
