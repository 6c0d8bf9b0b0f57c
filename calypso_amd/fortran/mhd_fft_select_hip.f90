!> mhd_fft_select_hip.f90 -- Fortran face of the GPU longitude FFTs, shaped
!! after the reference module MHD_FFT_selector
!! (MHD_src/sph_MHD/MHD_FFT_selector.F90):
!!
!!   init_MHD_FFT_select_hip(my_rank, l_truncation, nth, nr, nphi,
!!                           ncomp, ncomp_fwd, ncomp_bwd)
!!                                   ~ init_MHD_FFT_select        (:86-114)
!!   back_MHD_FFT_sel_from_recv_hip(ncomp_bwd, n_WR, WR, fld_rtp)
!!                                   ~ back_MHD_FFT_sel_from_recv (:219-249)
!!   fwd_MHD_FFT_sel_from_recv_hip(ncomp_fwd, n_WS, frc_rtp, WS)
!!                                   ~ fwd_MHD_FFT_sel_from_recv  (:185-215)
!!   finalize_MHD_FFT_select_hip     ~ finalize_MHD_FFT_select    (:118-145)
!!
!! The grid sizes the reference takes from m_spheric_parameter are arguments
!! of the init call.  WR / WS are the single-domain rtm arrays, component
!! fastest (stride ncomp_bwd / ncomp_fwd): what the reference's
!! self-communication leaves in its buffers.  fld_rtp / frc_rtp are
!! (nnod_rtp, ncomp) with the rtp node order k + nr (t-1) + nr nth (i-1).
!! A nonzero library status stops the run with the library's message.
      module mhd_fft_select_hip
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
      subroutine init_MHD_FFT_select_hip(my_rank, l_truncation,         &
     &          nth, nr, nphi, ncomp, ncomp_fwd, ncomp_bwd)
!
      integer(c_int), intent(in) :: my_rank
      integer(c_int), intent(in) :: l_truncation, nth, nr, nphi
      integer(c_int), intent(in) :: ncomp, ncomp_fwd, ncomp_bwd
!
      if(my_rank .eq. 0) write(*,*) 'Use HIP FFT'
      call check_status(calypso_init_MHD_FFT_select(l_truncation,       &
     &    nth, nr, nphi, ncomp, ncomp_fwd, ncomp_bwd))
!
      end subroutine init_MHD_FFT_select_hip
!
! -----------------------------------------------------------------------
!
      subroutine back_MHD_FFT_sel_from_recv_hip                         &
     &         (ncomp_bwd, n_WR, WR, fld_rtp)
!
      integer(c_int), intent(in) :: ncomp_bwd, n_WR
      real(c_double), intent(inout) :: WR(n_WR)
      real(c_double), intent(inout) :: fld_rtp(*)
!
      call check_status(calypso_back_MHD_FFT_sel_from_recv              &
     &    (ncomp_bwd, n_WR, WR, fld_rtp))
!
      end subroutine back_MHD_FFT_sel_from_recv_hip
!
! -----------------------------------------------------------------------
!
      subroutine fwd_MHD_FFT_sel_from_recv_hip                          &
     &         (ncomp_fwd, n_WS, frc_rtp, WS)
!
      integer(c_int), intent(in) :: ncomp_fwd, n_WS
      real(c_double), intent(inout) :: frc_rtp(*)
      real(c_double), intent(inout) :: WS(n_WS)
!
      call check_status(calypso_fwd_MHD_FFT_sel_from_recv               &
     &    (ncomp_fwd, n_WS, frc_rtp, WS))
!
      end subroutine fwd_MHD_FFT_sel_from_recv_hip
!
! -----------------------------------------------------------------------
!
      subroutine finalize_MHD_FFT_select_hip
!
      call check_status(calypso_finalize_MHD_FFT_select())
!
      end subroutine finalize_MHD_FFT_select_hip
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
      end module mhd_fft_select_hip
