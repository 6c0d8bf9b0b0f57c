!> legendre_trans_hip.f90 -- Fortran face of the GPU Legendre transforms,
!! shaped after the reference module legendre_transform_select
!! (SPH_SHELL_src/legendre_transform_select.f90):
!!
!!   set_legendre_trans_mode_hip(tranx_loop_ctl, id)  ~ set_legendre_trans_mode_ctl (:181-224)
!!   init_legendre_trans_hip(...)                     ~ sel_init_legendre_trans     (:228-265)
!!   leg_backward_trans_hip(ncomp, nvector, nscalar, n_WR, n_WS, WR, WS)
!!                                                    ~ sel_backward_legendre_trans (:309-371)
!!   leg_forward_trans_hip(ncomp, nvector, nscalar, n_WR, n_WS, WR, WS)
!!                                                    ~ sel_forward_legendre_trans  (:375-439)
!!   finalize_legendre_trans_hip                      ~ sel_finalize_legendre_trans (:269-304)
!!
!! WR / WS are the single-domain rlm and rtm arrays (component fastest), the
!! layout the reference's self-communication leaves in its buffers.  A nonzero
!! status from the library stops the run with the library's message, as the
!! reference does through calypso_MPI_abort (calypso_mpi.f90:114-126).
      module legendre_trans_hip
!
      use iso_c_binding
      use calypso_amd_capi
      implicit none
!
!>      integer flags of the GPU modes (beside the reference's 1..15, 99)
      integer(c_int), parameter :: iflag_leg_hip_fp64 =   101
      integer(c_int), parameter :: iflag_leg_hip_fp32 =   102
      integer(c_int), parameter :: iflag_leg_hip_bf16x3 = 103
      integer(c_int), parameter :: iflag_leg_hip_int8x6 = 104
!
      private :: check_status
!
      contains
!
! -----------------------------------------------------------------------
!
      subroutine set_legendre_trans_mode_hip(tranx_loop_ctl, id_leg)
!
      character(len = *), intent(in) :: tranx_loop_ctl
      integer(c_int), intent(inout) :: id_leg
      integer(c_int) :: mode
!
      mode = calypso_legendre_mode_id(trim(tranx_loop_ctl)//c_null_char)
      if(mode .gt. 0) id_leg = 100 + mode
!
      end subroutine set_legendre_trans_mode_hip
!
! -----------------------------------------------------------------------
!
      logical function is_legendre_trans_hip(id_leg)
      integer(c_int), intent(in) :: id_leg
      is_legendre_trans_hip = (id_leg .ge. iflag_leg_hip_fp64           &
     &                   .and. id_leg .le. iflag_leg_hip_int8x6)
      end function is_legendre_trans_hip
!
! -----------------------------------------------------------------------
!
      subroutine init_legendre_trans_hip(id_leg, l_truncation,          &
     &          nth, nr, radius_1d_rlm_r, g_colat_rtm, weight_rtm,      &
     &          ncomp, nvector, nscalar)
!
      integer(c_int), intent(in) :: id_leg, l_truncation, nth, nr
      real(c_double), intent(in) :: radius_1d_rlm_r(nr)
      real(c_double), intent(in) :: g_colat_rtm(nth), weight_rtm(nth)
      integer(c_int), intent(in) :: ncomp, nvector, nscalar
!
      call check_status(calypso_sel_init_legendre_trans(id_leg - 100,   &
     &    l_truncation, nth, nr, radius_1d_rlm_r, g_colat_rtm,          &
     &    weight_rtm, ncomp, nvector, nscalar))
!
      end subroutine init_legendre_trans_hip
!
! -----------------------------------------------------------------------
!
      subroutine leg_backward_trans_hip                                 &
     &         (ncomp, nvector, nscalar, n_WR, n_WS, WR, WS)
!
      integer(c_int), intent(in) :: ncomp, nvector, nscalar
      integer(c_int), intent(in) :: n_WR, n_WS
      real(c_double), intent(inout) :: WR(n_WR)
      real(c_double), intent(inout) :: WS(n_WS)
!
      call check_status(calypso_sel_backward_legendre_trans             &
     &    (ncomp, nvector, nscalar, n_WR, n_WS, WR, WS))
!
      end subroutine leg_backward_trans_hip
!
! -----------------------------------------------------------------------
!
      subroutine leg_forward_trans_hip                                  &
     &         (ncomp, nvector, nscalar, n_WR, n_WS, WR, WS)
!
      integer(c_int), intent(in) :: ncomp, nvector, nscalar
      integer(c_int), intent(in) :: n_WR, n_WS
      real(c_double), intent(inout) :: WR(n_WR)
      real(c_double), intent(inout) :: WS(n_WS)
!
      call check_status(calypso_sel_forward_legendre_trans              &
     &    (ncomp, nvector, nscalar, n_WR, n_WS, WR, WS))
!
      end subroutine leg_forward_trans_hip
!
! -----------------------------------------------------------------------
!
      subroutine finalize_legendre_trans_hip
!
      call check_status(calypso_sel_finalize_legendre_trans())
!
      end subroutine finalize_legendre_trans_hip
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
      end module legendre_trans_hip
