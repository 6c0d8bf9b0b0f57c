!> cal_nonlinear_sph_mhd_hip.f90 -- Fortran face of the GPU grid products,
!! after s_cal_nonlinear_sph_MHD (MHD_src/sph_MHD/cal_nonlinear_sph_MHD.f90:35-84),
!! which works on the module arrays fld_rtp / frc_rtp; here they are
!! arguments, in the transform order of the MHD step:
!!   fld_rtp(nnod_rtp, 13): velocity 1-3, vorticity 4-6, magnetic field 7-9,
!!                          current density 10-12, temperature 13
!!   frc_rtp(nnod_rtp, 19): w x u (coef_velo) 1-3, Coriolis 4-6 and its curl
!!                          7-9 (zero), J x B (coef_lor) 10-12, u x B
!!                          (coef_induct) 13-15, u T (coef_temp) 16-18,
!!                          Coriolis divergence 19 (zero)
!! Same operation order as the reference's cal_products_smp routines (bitwise
!! results).  A nonzero library status stops the run with its message.
      module cal_nonlinear_sph_mhd_hip
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
      subroutine s_cal_nonlinear_sph_MHD_hip(nnod_rtp, fld_rtp,         &
     &          frc_rtp, coef_velo, coef_lor, coef_induct, coef_temp)
!
      integer(c_int), intent(in) :: nnod_rtp
      real(c_double), intent(in) :: fld_rtp(nnod_rtp,13)
      real(c_double), intent(inout) :: frc_rtp(nnod_rtp,19)
      real(c_double), intent(in) :: coef_velo, coef_lor
      real(c_double), intent(in) :: coef_induct, coef_temp
      real(c_double) :: coef(4)
!
      coef = (/coef_velo, coef_lor, coef_induct, coef_temp/)
      call check_status(calypso_nonlinear_products                      &
     &    (int(nnod_rtp, c_long_long), fld_rtp, frc_rtp, coef))
!
      end subroutine s_cal_nonlinear_sph_MHD_hip
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
      end module cal_nonlinear_sph_mhd_hip
