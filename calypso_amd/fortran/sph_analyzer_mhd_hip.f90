!> sph_analyzer_mhd_hip.f90 -- Fortran face of the device-resident MHD step,
!! shaped after the reference module SPH_analyzer_MHD
!! (programs/SPH_MHD/SPH_analyzer_MHD.f90): SPH_initialize_MHD_hip takes
!! the initial fields a restart (or sph_initial_data_control) leaves in
!! d_rj, SPH_analyze_MHD_hip is one time step, and the output calls copy
!! back only what the output step needs.
!!
!! The state lives on the GPU between calls; the reference's d_rj is read
!! at initialisation and written only by sph_mhd_get_field_hip.
      module sph_analyzer_mhd_hip
!
      use iso_c_binding
      use calypso_amd_capi
      implicit none
!
      type(c_ptr), save, private :: mhd_ctx = c_null_ptr
!
      integer(c_int), parameter :: ihip_velo = 0
      integer(c_int), parameter :: ihip_vort = 3
      integer(c_int), parameter :: ihip_magne = 6
      integer(c_int), parameter :: ihip_current = 9
      integer(c_int), parameter :: ihip_temp = 12
!
      private :: check_status
!
      contains
!
! -----------------------------------------------------------------------
!
!>    Create the GPU state; with isig > 0 the dynamo_benchmark_1 initial
!!    fields (dynamo_benchmark_2's with ibench = 2) are set on the device,
!!    otherwise the caller provides velocity (nnod_rj,3), magnetic field
!!    (nnod_rj,3) and the full temperature.
!!    idx_gl_rj_j, when given, is the caller's local rj mode list
!!    (idx_gl_1d_rj_j(:,1) of m_spheric_parameter): its d_rj arrays, here
!!    and in sph_mhd_get_field_hip, are then in that mode order.
      subroutine SPH_initialize_MHD_hip(params, radius, isig,           &
     &          nnod_rj, velo, magne, temp, idx_gl_rj_j, ibench)
!
      type(calypso_mhd_params), intent(in) :: params
      real(c_double), intent(in) :: radius(*)
      integer(c_int), intent(in) :: isig, nnod_rj
      real(c_double), intent(in) :: velo(nnod_rj,3), magne(nnod_rj,3)
      real(c_double), intent(in) :: temp(nnod_rj)
      integer(c_int), intent(in), optional :: idx_gl_rj_j(:)
      integer(c_int), intent(in), optional :: ibench
      integer(c_int) :: nd
!
      call check_status(calypso_mhd_create(params, radius, c_null_ptr,  &
     &    mhd_ctx))
      if(present(idx_gl_rj_j)) then
        call check_status(calypso_mhd_set_rj_order(mhd_ctx,             &
     &      int(size(idx_gl_rj_j), c_int), idx_gl_rj_j))
      end if
      if(isig .gt. 0 .and. present(ibench)) then
        if(ibench .eq. 2) then
          call check_status(calypso_mhd_set_dynamobench2_ic(mhd_ctx,    &
     &        isig))
        else
          call check_status(calypso_mhd_set_dynamobench_ic(mhd_ctx,     &
     &        isig))
        end if
      else if(isig .gt. 0) then
        call check_status(calypso_mhd_set_dynamobench_ic(mhd_ctx, isig))
      else
        do nd = 1, 3
          call check_status(calypso_mhd_set_field(mhd_ctx,              &
     &        ihip_velo+nd-1, velo(1,nd)))
          call check_status(calypso_mhd_set_field(mhd_ctx,              &
     &        ihip_magne+nd-1, magne(1,nd)))
        end do
        call check_status(calypso_mhd_set_field(mhd_ctx, ihip_temp,     &
     &      temp))
      end if
      call check_status(calypso_mhd_initialize(mhd_ctx))
!
      end subroutine SPH_initialize_MHD_hip
!
! -----------------------------------------------------------------------
!
      subroutine SPH_analyze_MHD_hip(i_step, i_step_number,             &
     &          iflag_finish)
!
      integer(c_int), intent(in) :: i_step, i_step_number
      integer(c_int), intent(inout) :: iflag_finish
!
      call check_status(calypso_mhd_step(mhd_ctx, 1_c_int))
      if(i_step .ge. i_step_number .and. i_step_number.gt.0) then
        iflag_finish = 1
      end if
!
      end subroutine SPH_analyze_MHD_hip
!
! -----------------------------------------------------------------------
!
!>    The seven numbers of sph_pwr_volume_s.dat (K_ene_pol, K_ene_tor, K_ene,
!!    temperature, M_ene_pol, M_ene_tor, M_ene).
      subroutine sph_mhd_energies_hip(ene)
!
      real(c_double), intent(inout) :: ene(7)
!
      call check_status(calypso_mhd_energies(mhd_ctx, ene))
!
      end subroutine sph_mhd_energies_hip
!
! -----------------------------------------------------------------------
!
      subroutine sph_mhd_get_field_hip(icomp, nnod_rj, d_rj)
!
      integer(c_int), intent(in) :: icomp, nnod_rj
      real(c_double), intent(inout) :: d_rj(nnod_rj)
!
      call check_status(calypso_mhd_get_field(mhd_ctx, icomp, d_rj))
!
      end subroutine sph_mhd_get_field_hip
!
! -----------------------------------------------------------------------
!
!>    Fields on the physical grid at an output step
!!    (sph_back_trans_snapshot_MHD, sph_transforms_4_MHD.f90:200-224, without
!!    the pole values): nvec vectors (first d_rj component of each, e.g.
!!    ihip_velo) and nscl scalar components to fld_rtp(nnod_rtp, 3 nvec + nscl)
      subroutine sph_mhd_snapshot_rtp_hip(nvec, vec_comp, nscl,         &
     &          scl_comp, nnod_rtp, fld_rtp)
!
      integer(c_int), intent(in) :: nvec, nscl, nnod_rtp
      integer(c_int), intent(in) :: vec_comp(max(nvec,1))
      integer(c_int), intent(in) :: scl_comp(max(nscl,1))
      real(c_double), intent(inout) :: fld_rtp(nnod_rtp,3*nvec+nscl)
!
      call check_status(calypso_mhd_snapshot_rtp(mhd_ctx, nvec,         &
     &    vec_comp, nscl, scl_comp, fld_rtp))
!
      end subroutine sph_mhd_snapshot_rtp_hip
!
! -----------------------------------------------------------------------
!
      subroutine SPH_finalize_MHD_hip
!
      call check_status(calypso_mhd_destroy(mhd_ctx))
      mhd_ctx = c_null_ptr
!
      end subroutine SPH_finalize_MHD_hip
!
! -----------------------------------------------------------------------
!
      subroutine check_status(ierr)
!
      integer(c_int), intent(in) :: ierr
      type(c_ptr) :: cmsg
      character(kind=c_char), pointer :: msg(:)
      integer :: n
      interface
        integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
          import :: c_size_t, c_ptr
          type(c_ptr), value :: s
        end function c_strlen
      end interface
!
      if(ierr .eq. 0) return
      cmsg = calypso_last_error()
      n = int(c_strlen(cmsg))
      call c_f_pointer(cmsg, msg, [n])
      write(*,*) 'calypso_amd MHD step failed: ', msg(1:n)
      error stop 1
!
      end subroutine check_status
!
! -----------------------------------------------------------------------
!
      end module sph_analyzer_mhd_hip
