!> calypso_amd_capi.f90 -- ISO_C_BINDING interfaces of the C ABI declared in
!! include/calypso_amd.h (libcalypso_amd.so).  Integers are c_int (Calypso's
!! kint = 4), reals c_double (kreal = 8), arrays passed by address.
      module calypso_amd_capi
!
      use iso_c_binding
      implicit none
!
      integer(c_int), parameter :: CALYPSO_LEG_FP64 =   1
      integer(c_int), parameter :: CALYPSO_LEG_FP32 =   2
      integer(c_int), parameter :: CALYPSO_LEG_BF16X3 = 3
!
      interface
!
        function calypso_last_error() bind(C, name="calypso_last_error")
          import :: c_ptr
          type(c_ptr) :: calypso_last_error
        end function calypso_last_error
!
        integer(c_int) function calypso_legendre_mode_id(name)          &
     &      bind(C, name="calypso_legendre_mode_id")
          import :: c_int, c_char
          character(kind=c_char), dimension(*), intent(in) :: name
        end function calypso_legendre_mode_id
!
        integer(c_int) function calypso_sel_init_legendre_trans(mode,   &
     &      l_truncation, nth, nr, radius, colat, weight,               &
     &      ncomp, nvector, nscalar)                                    &
     &      bind(C, name="calypso_sel_init_legendre_trans")
          import :: c_int, c_double
          integer(c_int), value :: mode, l_truncation, nth, nr
          real(c_double), intent(in) :: radius(*), colat(*), weight(*)
          integer(c_int), value :: ncomp, nvector, nscalar
        end function calypso_sel_init_legendre_trans
!
        integer(c_int) function calypso_sel_backward_legendre_trans(    &
     &      ncomp, nvector, nscalar, n_WR, n_WS, WR, WS)                &
     &      bind(C, name="calypso_sel_backward_legendre_trans")
          import :: c_int, c_double
          integer(c_int), value :: ncomp, nvector, nscalar, n_WR, n_WS
          real(c_double), intent(in) :: WR(*)
          real(c_double), intent(inout) :: WS(*)
        end function calypso_sel_backward_legendre_trans
!
        integer(c_int) function calypso_sel_forward_legendre_trans(     &
     &      ncomp, nvector, nscalar, n_WR, n_WS, WR, WS)                &
     &      bind(C, name="calypso_sel_forward_legendre_trans")
          import :: c_int, c_double
          integer(c_int), value :: ncomp, nvector, nscalar, n_WR, n_WS
          real(c_double), intent(in) :: WR(*)
          real(c_double), intent(inout) :: WS(*)
        end function calypso_sel_forward_legendre_trans
!
        integer(c_int) function calypso_sel_finalize_legendre_trans()   &
     &      bind(C, name="calypso_sel_finalize_legendre_trans")
          import :: c_int
        end function calypso_sel_finalize_legendre_trans
!
      end interface
!
      end module calypso_amd_capi
