! ref_derivs_harness.f90 -- drives the reference's own derivs module.
!
! TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile against the unchanged
! reference sources parameters.f90 and derivs.f90 (compiled where they lie),
! with a grid fixed by the parameters.in this recipe writes.  Output goes to
! oracle/_ref/.  It pins the C oracle's stencils and boundary index maps to
! the reference: deriv_xx/yy/zz (derivs.f90:146-272) summed as laplacian
! does (variables.f90:68), and deriv_x (derivs.f90:25-64).
!
! Usage: ref_derivs <bcs> <order> <dx> <dy> <dz> <in.bin> <out.bin>
!   in.bin : nx*ny*nz complex(pr), x fastest (no ghosts)
!   out.bin: laplacian then deriv_x, same layout
! Ghost planes are filled as the one-rank halo exchange fills them
! (variables.f90:203-375): the wrapped planes for bcs=1, and NaN for bcs=2,
! where the reflective index maps must never read them.
program ref_derivs_harness
  use parameters
  use derivs
  implicit none

  complex (pr), allocatable :: f(:,:,:), dxx(:,:,:), dyy(:,:,:), dzz(:,:,:), &
    lap(:,:,:), fx(:,:,:), buf(:,:,:)
  character(len=256) :: arg, fin, fout
  real (pr) :: nan_v
  integer :: u

  call get_command_argument(1, arg); read (arg, *) bcs
  call get_command_argument(2, arg); read (arg, *) order
  call get_command_argument(3, arg); read (arg, *) dx
  call get_command_argument(4, arg); read (arg, *) dy
  call get_command_argument(5, arg); read (arg, *) dz
  call get_command_argument(6, fin)
  call get_command_argument(7, fout)

  js = 0; je = ny-1; ks = 0; ke = nz-1
  dx2 = dx**2; dy2 = dy**2; dz2 = dz**2

  allocate(f(0:nx1,js-2:je+2,ks-2:ke+2))
  allocate(dxx(0:nx1,js:je,ks:ke), dyy(0:nx1,js:je,ks:ke), dzz(0:nx1,js:je,ks:ke))
  allocate(lap(0:nx1,js:je,ks:ke), fx(0:nx1,js:je,ks:ke))

  nan_v = 0.0_pr
  nan_v = nan_v / nan_v
  f = cmplx(nan_v, nan_v, pr)

  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  allocate(buf(0:nx1,js:je,ks:ke))
  read (u) buf
  close (u)
  f(:,js:je,ks:ke) = buf

  if (bcs == 1) then
    f(:,js:je,ks-2) = f(:,js:je,ke-1)
    f(:,js:je,ks-1) = f(:,js:je,ke)
    f(:,js:je,ke+1) = f(:,js:je,ks)
    f(:,js:je,ke+2) = f(:,js:je,ks+1)
    f(:,js-2,:) = f(:,je-1,:)
    f(:,js-1,:) = f(:,je,:)
    f(:,je+1,:) = f(:,js,:)
    f(:,je+2,:) = f(:,js+1,:)
  end if

  call deriv_xx(f, dxx)
  call deriv_yy(f, dyy)
  call deriv_zz(f, dzz)
  lap = dxx + dyy + dzz
  call deriv_x(f, fx)

  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) lap
  write (u) fx
  close (u)
end program ref_derivs_harness
