! Calls libconv_amd.so from Fortran through the interface module
! (include/libconv_amd.f90) and prints the results for tests/test_fortran.py:
!   1. d_s0s0_1d_sym8_md (generated C name) along dim 1 of a 12 x 10 x 8
!      array, periodic;
!   2. s0s0_1d_sym8_md_d_ (libconv.f90 spelling) on the same data, which
!      must give the same bits;
!   3. d_sym8_d2_dotin (3D Kinetic with energies), free grow;
!   4. an invalid call (bc 7), which records an error instead of running.
! Inputs: x(i) = sin(0.37 i) + 0.25 cos(1.3 i), i = 1..size (no RNG).
program lc_fortran_check
  use iso_c_binding
  use libconv_amd
  implicit none
  integer(c_int), parameter :: n1 = 12, n2 = 10, n3 = 8
  real(c_double) :: x(n1 * n2 * n3), y(n1 * n2 * n3), y2(n1 * n2 * n3)
  real(c_double) :: xk(n1 * n2 * n3), yk((n1 + 28) * (n2 + 28) * (n3 + 28)), w(1), dots(3), a3(3)
  integer(c_int) :: n(3), bc3(3), i, code
  n = [n1, n2, n3]
  do i = 1, size(x)
    x(i) = sin(0.37d0 * i) + 0.25d0 * cos(1.3d0 * i)
  end do
  y = 0
  call d_s0s0_1d_sym8_md(3, 1, n, LC_BC_PERIODIC, n, n, 1, x, y, 1.0d0, 0.0d0, 0.0d0)
  if (lc_last_error() /= 0) stop 2
  y2 = 0
  call s0s0_1d_sym8_md_d(3, 1, n, LC_BC_PERIODIC, n, n, 1, 1.0d0, 0.0d0, 0.0d0, x, y2)
  if (lc_last_error() /= 0) stop 3
  write(*, '(a, i0)') 'same_bits ', merge(1, 0, all(y == y2))
  write(*, '(a)') 'mf'
  write(*, '(4es25.17)') y
  xk = x
  yk = 0
  bc3 = LC_BC_GROW
  a3 = [-0.5d0 / 0.3d0**2, -0.5d0 / 0.35d0**2, -0.5d0 / 0.4d0**2]
  call d_sym8_d2_dotin(3, n, bc3, xk, yk, w, w, a3, 0.0d0, 0.0d0, dots)
  if (lc_last_error() /= 0) stop 4
  write(*, '(a)') 'kin'
  write(*, '(4es25.17)') yk
  write(*, '(a)') 'dots'
  write(*, '(3es25.17)') dots
  call d_s0s0_1d_sym8_md(3, 1, n, 7, n, n, 1, x, y, 1.0d0, 0.0d0, 0.0d0)
  code = lc_last_error()
  write(*, '(a, i0)') 'bad_bc_error ', code
end program lc_fortran_check
