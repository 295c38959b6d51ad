! dropin_driver.f90 -- drives the drop-in module solve the way gpe.f90 does.
!
! TEST HARNESS.  Built by oracle/Makefile (target dropin) against the
! reference's unchanged parameters.f90 and error.f90 and the drop-in
! gpe_amd/fortran/solve.f90.  It reads the reference's own namelists
! run_params and eqn_params (parameters.f90:40-56) from a run.in-style file,
! forms the grid as get_grid does (ic.f90:43-66), reads the initial state in
! the restart format of end_state (io.f90:843-853), and runs the time loop
! of gpe.f90:136-280 (copy psi%new into psi%old, call solver, p = p+1)
! while t+im_t <= end_time.  The final state is written in the same format.
!
! Usage: dropin_driver <run.in> <state_in> <state_out>
program dropin_driver
  use parameters
  use solve
  implicit none

  complex (pr), allocatable :: old(:,:,:), new(:,:,:)
  character(len=512) :: fin_nml, fin, fout
  integer :: u, nx_in, ny_in, nz_in, p_in
  real (pr) :: t_in, im_t_in
  complex (pr) :: dt_in

  call MPI_INIT(ierr)
  call get_command_argument(1, fin_nml)
  call get_command_argument(2, fin)
  call get_command_argument(3, fout)

  open (newunit=u, file=trim(fin_nml), status='old')
  read (u, nml=run_params)
  read (u, nml=eqn_params)
  close (u)

  myrank = 0
  js = 0; je = ny-1; ks = 0; ke = nz-1
  save_rate = 1
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

  old = 0.0_pr
  do while (t+im_t <= end_time)
    old(:,js:je,ks:ke) = new
    call solver(old, new)
    p = p+1
  end do

  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) nx, ny, nz, p
  write (u) t, im_t
  write (u) dt
  write (u) new
  close (u)
  call MPI_FINALIZE(ierr)
end program dropin_driver
