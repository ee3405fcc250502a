! libconv_amd.f90 -- Fortran interface module of libconv_amd.so.
!
! The library exports three families of names (include/libconv.h):
!   * the C names BOAST generates (d_s0s0_1d_sym8_md, d_sym8_d2_dotin, ...):
!     integers and reals by value, arrays by address.  This module declares
!     them with bind(C), so Fortran callers of the generated procedures
!     (BigDFT, BigDFT/TestBuild.rb:12-28) call them with type checking;
!   * the libconv.f90 spellings (BigDFT/libconv.f90:303-456): every argument
!     by reference, lower case with a trailing underscore.  Existing calls
!     compile without this module; it declares them too, with their
!     argument kinds;
!   * the device-pointer stream API (lc_conv_md, lc_conv_1ds, lc_wavelet_md)
!     and the local Hamiltonian, which take type(c_ptr) device addresses
!     and return an error code (0 = success).
! Arrays are Fortran-ordered (dim 0 fastest), as the reference passes them.
module libconv_amd
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: LC_BC_PERIODIC = 0, LC_BC_GROW = 1, LC_BC_SHRINK = -1, LC_BC_FREE = 2
  integer(c_int), parameter :: LC_FIL_MD = 0, LC_FIL_IMD = 1, LC_FIL_D2 = 2, LC_FIL_D1 = 3
  integer(c_int), parameter :: LC_WAV_SYM8 = 0, LC_WAV_SYMICOMB8 = 1

  interface
    ! ---------------------------------------------------------- errors
    integer(c_int) function lc_last_error() bind(C, name='lc_last_error')
      import :: c_int
    end function
    type(c_ptr) function lc_last_error_message() bind(C, name='lc_last_error_message')
      import :: c_ptr
    end function
    type(c_ptr) function lc_version() bind(C, name='lc_version')
      import :: c_ptr
    end function

    ! ------------------------------------------ generated 1D procedures
    ! y = a fil(x) + a_x x + a_y y along dim idim (0-based) of a d-dim array
    subroutine d_s0s0_1d_sym8_md(d, idim, n, bc, nx, ny, narr, x, y, a, a_x, a_y) &
        bind(C, name='d_s0s0_1d_sym8_md')
      import :: c_int, c_double
      integer(c_int), value :: d, idim, bc, narr
      integer(c_int), intent(in) :: n(*), nx(*), ny(*)
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: y(*)
      real(c_double), value :: a, a_x, a_y
    end subroutine
    subroutine s_s0s0_1d_sym8_md(d, idim, n, bc, nx, ny, narr, x, y, a, a_x, a_y) &
        bind(C, name='s_s0s0_1d_sym8_md')
      import :: c_int, c_float
      integer(c_int), value :: d, idim, bc, narr
      integer(c_int), intent(in) :: n(*), nx(*), ny(*)
      real(c_float), intent(in) :: x(*)
      real(c_float), intent(inout) :: y(*)
      real(c_float), value :: a, a_x, a_y
    end subroutine
    subroutine d_s0s0_1d_sym8_imd(d, idim, n, bc, nx, ny, narr, x, y, a, a_x, a_y) &
        bind(C, name='d_s0s0_1d_sym8_imd')
      import :: c_int, c_double
      integer(c_int), value :: d, idim, bc, narr
      integer(c_int), intent(in) :: n(*), nx(*), ny(*)
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: y(*)
      real(c_double), value :: a, a_x, a_y
    end subroutine
    subroutine d_s0s0_1d_sym8_d2(d, idim, n, bc, nx, ny, narr, x, y, a, a_x, a_y) &
        bind(C, name='d_s0s0_1d_sym8_d2')
      import :: c_int, c_double
      integer(c_int), value :: d, idim, bc, narr
      integer(c_int), intent(in) :: n(*), nx(*), ny(*)
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: y(*)
      real(c_double), value :: a, a_x, a_y
    end subroutine
    subroutine d_s0s0_1d_sym8_d2_dotin(d, idim, n, bc, nx, ny, narr, x, y, a, a_x, a_y, dot_in) &
        bind(C, name='d_s0s0_1d_sym8_d2_dotin')
      import :: c_int, c_double
      integer(c_int), value :: d, idim, bc, narr
      integer(c_int), intent(in) :: n(*), nx(*), ny(*)
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: y(*)
      real(c_double), value :: a, a_x, a_y
      real(c_double), intent(out) :: dot_in
    end subroutine
    subroutine d_s0s1_1d_sym8(d, idim, n, bc, nx, ny, narr, x, y, a, a_y) &
        bind(C, name='d_s0s1_1d_sym8')
      import :: c_int, c_double
      integer(c_int), value :: d, idim, bc, narr
      integer(c_int), intent(in) :: n(*), nx(*), ny(*)
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: y(*)
      real(c_double), value :: a, a_y
    end subroutine
    subroutine d_s1s0_1d_sym8(d, idim, n, bc, nx, ny, narr, x, y, a, a_y) &
        bind(C, name='d_s1s0_1d_sym8')
      import :: c_int, c_double
      integer(c_int), value :: d, idim, bc, narr
      integer(c_int), intent(in) :: n(*), nx(*), ny(*)
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: y(*)
      real(c_double), value :: a, a_y
    end subroutine
    subroutine d_s0s1_1d_symicomb8(d, idim, n, bc, nx, ny, narr, x, y, a, a_y) &
        bind(C, name='d_s0s1_1d_symicomb8')
      import :: c_int, c_double
      integer(c_int), value :: d, idim, bc, narr
      integer(c_int), intent(in) :: n(*), nx(*), ny(*)
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: y(*)
      real(c_double), value :: a, a_y
    end subroutine
    subroutine d_s1s0_1d_symicomb8(d, idim, n, bc, nx, ny, narr, x, y, a, a_y) &
        bind(C, name='d_s1s0_1d_symicomb8')
      import :: c_int, c_double
      integer(c_int), value :: d, idim, bc, narr
      integer(c_int), intent(in) :: n(*), nx(*), ny(*)
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: y(*)
      real(c_double), value :: a, a_y
    end subroutine

    ! ------------------------------------------------- 3D composers
    subroutine d_sym8_md_ld_narr(d, n, bc, nx, ny, narr, x, y, w1, w2) bind(C, name='d_sym8_md_ld_narr')
      import :: c_int, c_double
      integer(c_int), value :: d, narr
      integer(c_int), intent(in) :: n(*), bc(*), nx(*), ny(*)
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: y(*), w1(*), w2(*)
    end subroutine
    subroutine d_sym8_d2_dotin(d, n, bc, x, y, w1, w2, a, a_x, a_y, dot_in) bind(C, name='d_sym8_d2_dotin')
      import :: c_int, c_double
      integer(c_int), value :: d
      integer(c_int), intent(in) :: n(*), bc(*)
      real(c_double), intent(in) :: x(*), a(*)
      real(c_double), intent(inout) :: y(*), w1(*), w2(*)
      real(c_double), value :: a_x, a_y
      real(c_double), intent(out) :: dot_in(*)
    end subroutine
    subroutine d_dwt_sym8(d, n, bc, x, y, w1, w2) bind(C, name='d_dwt_sym8')
      import :: c_int, c_double
      integer(c_int), value :: d
      integer(c_int), intent(in) :: n(*), bc(*)
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: y(*), w1(*), w2(*)
    end subroutine
    subroutine d_idwt_sym8(d, n, bc, x, y, w1, w2) bind(C, name='d_idwt_sym8')
      import :: c_int, c_double
      integer(c_int), value :: d
      integer(c_int), intent(in) :: n(*), bc(*)
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(inout) :: y(*), w1(*), w2(*)
    end subroutine

    ! ------------------------------- libconv.f90 spellings (by reference)
    subroutine s0s0_1d_sym8_md_d(d, idim, n, bc_id, nx, ny, narr, a, a_x, a_y, x, y) &
        bind(C, name='s0s0_1d_sym8_md_d_')
      import :: c_int, c_double
      integer(c_int), intent(in) :: d, idim, n(*), bc_id, nx(*), ny(*), narr
      real(c_double), intent(in) :: a, a_x, a_y, x(*)
      real(c_double), intent(inout) :: y(*)
    end subroutine
    subroutine s0s0_1ds_dot_sym8_d2_d(d, n, bc, nx, ny, narr, a, a_x, a_y, x, y, dot_in, work) &
        bind(C, name='s0s0_1ds_dot_sym8_d2_d_')
      import :: c_int, c_double
      integer(c_int), intent(in) :: d, n(*), bc(*), nx(*), ny(*), narr
      real(c_double), intent(in) :: a(*), a_x, a_y, x(*)
      real(c_double), intent(inout) :: y(*), work(*)
      real(c_double), intent(out) :: dot_in(*)
    end subroutine

    ! ------------------------------------- stream API (device memory)
    integer(c_int) function lc_conv_md(prec, fil, d, n, bc, nx, ny, narr, x, y, a, a_x, a_y, stream) &
        bind(C, name='lc_conv_md')
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: prec, fil, d, narr
      integer(c_int), intent(in) :: n(*), bc(*), nx(*), ny(*)
      type(c_ptr), value :: x, y, stream
      real(c_double), value :: a, a_x, a_y
    end function
    integer(c_int) function lc_conv_1ds(prec, fil, d, n, bc, nx, ny, narr, a, a_x, a_y, x, y, dot, stream) &
        bind(C, name='lc_conv_1ds')
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: prec, fil, d, narr
      integer(c_int), intent(in) :: n(*), bc(*), nx(*), ny(*)
      real(c_double), intent(in) :: a(*)
      real(c_double), value :: a_x, a_y
      type(c_ptr), value :: x, y, dot, stream
    end function
    integer(c_int) function lc_wavelet_md(prec, wav, inverse, d, n, bc, nx, ny, narr, x, y, a, a_y, stream) &
        bind(C, name='lc_wavelet_md')
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: prec, wav, inverse, d, narr
      integer(c_int), intent(in) :: n(*), bc(*), nx(*), ny(*)
      type(c_ptr), value :: x, y, stream
      real(c_double), value :: a, a_y
    end function
    integer(c_long) function lc_local_hamiltonian_work(n) bind(C, name='lc_local_hamiltonian_work')
      import :: c_int, c_long
      integer(c_int), intent(in) :: n(*)
    end function
    integer(c_int) function lc_local_hamiltonian(prec, n, a, psig, hpsig, pot, rho, work, ekin, stream) &
        bind(C, name='lc_local_hamiltonian')
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: prec
      integer(c_int), intent(in) :: n(*)
      real(c_double), intent(in) :: a(*)
      type(c_ptr), value :: psig, hpsig, pot, rho, work, ekin, stream
    end function
  end interface

contains

  ! text of the last error of this host thread ('' when none)
  function lc_error_text() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: s(:)
    integer :: k
    p = lc_last_error_message()
    msg = ''
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [4096])
    k = 0
    do while (k < 4096)
      if (s(k + 1) == c_null_char) exit
      k = k + 1
    end do
    allocate(character(len=k) :: msg)
    msg = transfer(s(1:k), msg)
  end function

end module libconv_amd
