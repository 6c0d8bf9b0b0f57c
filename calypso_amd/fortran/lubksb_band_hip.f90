!> lubksb_band_hip.f90 -- Fortran face of the GPU band solves, with the
!! argument lists of the reference module lubksb_357band_mul
!! (SOLVER_src/direct_solver/lubksb_357band_mul.f90:73-195):
!!
!!   lubksb_3band_mul_hip(Msmp, Msmp_stack, mcomp, n, band_lu, i_pivot, x)
!!                                          ~ lubksb_3band_mul
!!   lubksb_5band_mul_hip(Msmp, Msmp_stack, mcomp, n, band_lu, i_pivot, x)
!!                                          ~ lubksb_5band_mul
!!
!! band_lu(2 nb - 1, n, mcomp) and i_pivot(n, mcomp) from the reference's
!! ludcmp_3band / ludcmp_band; x(mcomp, n) is solved in place for the modes
!! Msmp_stack(0)+1 .. Msmp_stack(Msmp), with the reference's row interchanges
!! and operation order (bitwise the reference's results).  A nonzero library
!! status (a pivot outside 1..n, a bad stack) stops the run with its message.
      module lubksb_band_hip
!
      use iso_c_binding
      use calypso_amd_capi
      implicit none
!
      private :: check_status
!
      contains
!
! -----------------------------------------------------------------------
!
      subroutine lubksb_3band_mul_hip(Msmp, Msmp_stack, mcomp, n,       &
     &          band_lu, i_pivot, x)
!
      integer(c_int), intent(in) :: Msmp, mcomp, n
      integer(c_int), intent(in) :: Msmp_stack(0:Msmp)
      integer(c_int), intent(in) :: i_pivot(n,mcomp)
      real(c_double), intent(in) :: band_lu(5,n,mcomp)
      real(c_double), intent(inout) :: x(mcomp,n)
!
      call check_status(calypso_lubksb_3band_mul(Msmp, Msmp_stack,      &
     &    mcomp, n, band_lu, i_pivot, x))
!
      end subroutine lubksb_3band_mul_hip
!
! -----------------------------------------------------------------------
!
      subroutine lubksb_5band_mul_hip(Msmp, Msmp_stack, mcomp, n,       &
     &          band_lu, i_pivot, x)
!
      integer(c_int), intent(in) :: Msmp, mcomp, n
      integer(c_int), intent(in) :: Msmp_stack(0:Msmp)
      integer(c_int), intent(in) :: i_pivot(n,mcomp)
      real(c_double), intent(in) :: band_lu(9,n,mcomp)
      real(c_double), intent(inout) :: x(mcomp,n)
!
      call check_status(calypso_lubksb_5band_mul(Msmp, Msmp_stack,      &
     &    mcomp, n, band_lu, i_pivot, x))
!
      end subroutine lubksb_5band_mul_hip
!
! -----------------------------------------------------------------------
!
      subroutine check_status(ierr)
!
      integer(c_int), intent(in) :: ierr
      character(kind=c_char), pointer :: msg(:)
      character(len=512) :: text
      integer :: i
!
      if(ierr .eq. 0) return
      call c_f_pointer(calypso_last_error(), msg, [512])
      text = ' '
      do i = 1, 512
        if(msg(i) .eq. c_null_char) exit
        text(i:i) = msg(i)
      end do
      write(*,*) 'calypso_amd error: ', trim(text)
      error stop 1
!
      end subroutine check_status
!
! -----------------------------------------------------------------------
!
      end module lubksb_band_hip
