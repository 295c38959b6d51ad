! fortran_e_fixture.f90 -- writes values with the edit descriptor the
! reference's zeros files use ('(3e17.9)') and one list-directed string,
! for tests/test_zeros_oracle.py.  Build and run (from tests/golden):
!   flang -O0 fortran_e_fixture.f90 -o /tmp/fe && /tmp/fe > fortran_e_fixture.txt
program fortran_e_fixture
  implicit none
  real(8) :: v(12)
  real(4) :: s(3)
  v = [1.0d0, -12.345678912345d0, 0.0d0, 3.2d-120, 9.9999999996d0, -0.5d0, &
       123456.789d0, 1.0d-5, -7.77777777777d33, 2.5d0, 0.15d0, 1.0d100]
  write (*, *) "# i,j,k --> i+1,j,k"
  write (*, '(3e17.9)') v(1), v(2), v(3)
  write (*, '(3e17.9)') v(4), v(5), v(6)
  write (*, '(3e17.9)') v(7), v(8), v(9)
  write (*, '(3e17.9)') v(10), v(11), v(12)
  s = [0.1e0, -3.14159274e0, 6.0e-8]
  write (*, '(3e17.9)') s(1), s(2), s(3)
end program fortran_e_fixture
