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
!>    calypso_mhd_params of include/calypso_amd.h (same field order)
      type, bind(C) :: calypso_mhd_params
        integer(c_int) :: l_truncation, nth, nphi, nr
        integer(c_int) :: legendre_mode
        real(c_double) :: dt
        real(c_double) :: coef_imp_v, coef_imp_t, coef_imp_b
        real(c_double) :: coef_velo, coef_d_velo, coef_buo, coef_cor
        real(c_double) :: coef_lor
        real(c_double) :: coef_temp, coef_d_temp
        real(c_double) :: coef_magne, coef_d_magne, coef_induct
        real(c_double) :: depth_high_t, depth_low_t
        real(c_double) :: high_temp, low_temp
        integer(c_int) :: bc_velo_icb, bc_velo_cmb
        integer(c_int) :: kr_icb, kr_cmb
        integer(c_int) :: bc_magne_icb
        real(c_double) :: coef_press
      end type calypso_mhd_params
      integer(c_int), parameter :: CALYPSO_VELO_RIGID =          0
      integer(c_int), parameter :: CALYPSO_VELO_FREE_SLIP =      1
      integer(c_int), parameter :: CALYPSO_VELO_ROT_INNER_CORE = 2
      integer(c_int), parameter :: CALYPSO_MAGNE_INSULATOR =     0
      integer(c_int), parameter :: CALYPSO_MAGNE_TO_CENTER =     1
!
!>    calypso_scalar_bc of include/calypso_amd.h (same field order)
      integer(c_int), parameter :: CALYPSO_BC_FIXED_VALUE = 0
      integer(c_int), parameter :: CALYPSO_BC_FIXED_FLUX =  1
      type, bind(C) :: calypso_scalar_bc
        real(c_double) :: coef_d, coef_imp
        integer(c_int) :: bc_icb, bc_cmb
        type(c_ptr) :: name
      end type calypso_scalar_bc
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
        integer(c_int) function calypso_init_MHD_FFT_select(            &
     &      l_truncation, nth, nr, nphi, ncomp, ncomp_fwd, ncomp_bwd)   &
     &      bind(C, name="calypso_init_MHD_FFT_select")
          import :: c_int
          integer(c_int), value :: l_truncation, nth, nr, nphi
          integer(c_int), value :: ncomp, ncomp_fwd, ncomp_bwd
        end function calypso_init_MHD_FFT_select
!
        integer(c_int) function calypso_back_MHD_FFT_sel_from_recv(     &
     &      ncomp_bwd, n_WR, WR, fld_rtp)                               &
     &      bind(C, name="calypso_back_MHD_FFT_sel_from_recv")
          import :: c_int, c_double
          integer(c_int), value :: ncomp_bwd, n_WR
          real(c_double), intent(in) :: WR(*)
          real(c_double), intent(inout) :: fld_rtp(*)
        end function calypso_back_MHD_FFT_sel_from_recv
!
        integer(c_int) function calypso_fwd_MHD_FFT_sel_from_recv(      &
     &      ncomp_fwd, n_WS, frc_rtp, WS)                               &
     &      bind(C, name="calypso_fwd_MHD_FFT_sel_from_recv")
          import :: c_int, c_double
          integer(c_int), value :: ncomp_fwd, n_WS
          real(c_double), intent(in) :: frc_rtp(*)
          real(c_double), intent(inout) :: WS(*)
        end function calypso_fwd_MHD_FFT_sel_from_recv
!
        integer(c_int) function calypso_nonlinear_products(nnod,        &
     &      fld_rtp, frc_rtp, coef)                                     &
     &      bind(C, name="calypso_nonlinear_products")
          import :: c_int, c_double, c_long_long
          integer(c_long_long), value :: nnod
          real(c_double), intent(in) :: fld_rtp(*), coef(4)
          real(c_double), intent(inout) :: frc_rtp(*)
        end function calypso_nonlinear_products
!
        integer(c_int) function calypso_finalize_MHD_FFT_select()       &
     &      bind(C, name="calypso_finalize_MHD_FFT_select")
          import :: c_int
        end function calypso_finalize_MHD_FFT_select
!
        integer(c_int) function calypso_lubksb_3band_mul(Msmp,          &
     &      Msmp_stack, mcomp, n, band_lu, i_pivot, x)                  &
     &      bind(C, name="calypso_lubksb_3band_mul")
          import :: c_int, c_double
          integer(c_int), value :: Msmp, mcomp, n
          integer(c_int), intent(in) :: Msmp_stack(*), i_pivot(*)
          real(c_double), intent(in) :: band_lu(*)
          real(c_double), intent(inout) :: x(*)
        end function calypso_lubksb_3band_mul
!
        integer(c_int) function calypso_lubksb_5band_mul(Msmp,          &
     &      Msmp_stack, mcomp, n, band_lu, i_pivot, x)                  &
     &      bind(C, name="calypso_lubksb_5band_mul")
          import :: c_int, c_double
          integer(c_int), value :: Msmp, mcomp, n
          integer(c_int), intent(in) :: Msmp_stack(*), i_pivot(*)
          real(c_double), intent(in) :: band_lu(*)
          real(c_double), intent(inout) :: x(*)
        end function calypso_lubksb_5band_mul
!
        integer(c_int) function calypso_mhd_create(params, radius,      &
     &      stream, mhd) bind(C, name="calypso_mhd_create")
          import :: c_int, c_double, c_ptr, calypso_mhd_params
          type(calypso_mhd_params), intent(in) :: params
          real(c_double), intent(in) :: radius(*)
          type(c_ptr), value :: stream
          type(c_ptr), intent(out) :: mhd
        end function calypso_mhd_create
!
        integer(c_int) function calypso_mhd_destroy(mhd)                &
     &      bind(C, name="calypso_mhd_destroy")
          import :: c_int, c_ptr
          type(c_ptr), value :: mhd
        end function calypso_mhd_destroy
!
        integer(c_int) function calypso_mhd_set_field(mhd, comp, d)     &
     &      bind(C, name="calypso_mhd_set_field")
          import :: c_int, c_double, c_ptr
          type(c_ptr), value :: mhd
          integer(c_int), value :: comp
          real(c_double), intent(in) :: d(*)
        end function calypso_mhd_set_field
!
        integer(c_int) function calypso_mhd_get_field(mhd, comp, d)     &
     &      bind(C, name="calypso_mhd_get_field")
          import :: c_int, c_double, c_ptr
          type(c_ptr), value :: mhd
          integer(c_int), value :: comp
          real(c_double), intent(inout) :: d(*)
        end function calypso_mhd_get_field
!
        integer(c_int) function calypso_mhd_set_rj_order(mhd, jmax, jgl)  &
     &      bind(C, name="calypso_mhd_set_rj_order")
          import :: c_int, c_ptr
          type(c_ptr), value :: mhd
          integer(c_int), value :: jmax
          integer(c_int), intent(in) :: jgl(*)
        end function calypso_mhd_set_rj_order
!
        integer(c_int) function calypso_mhd_set_dynamobench_ic(mhd,     &
     &      isig) bind(C, name="calypso_mhd_set_dynamobench_ic")
          import :: c_int, c_ptr
          type(c_ptr), value :: mhd
          integer(c_int), value :: isig
        end function calypso_mhd_set_dynamobench_ic
        integer(c_int) function calypso_mhd_set_dynamobench2_ic(mhd,     &
     &      isig) bind(C, name="calypso_mhd_set_dynamobench2_ic")
          import :: c_int, c_ptr
          type(c_ptr), value :: mhd
          integer(c_int), value :: isig
        end function calypso_mhd_set_dynamobench2_ic
!
        integer(c_int) function calypso_mhd_initialize(mhd)             &
     &      bind(C, name="calypso_mhd_initialize")
          import :: c_int, c_ptr
          type(c_ptr), value :: mhd
        end function calypso_mhd_initialize
!
        integer(c_int) function calypso_mhd_step(mhd, nsteps)           &
     &      bind(C, name="calypso_mhd_step")
          import :: c_int, c_ptr
          type(c_ptr), value :: mhd
          integer(c_int), value :: nsteps
        end function calypso_mhd_step
!
        integer(c_int) function calypso_mhd_energies(mhd, ene)          &
     &      bind(C, name="calypso_mhd_energies")
          import :: c_int, c_double, c_ptr
          type(c_ptr), value :: mhd
          real(c_double), intent(inout) :: ene(7)
        end function calypso_mhd_energies
!
        integer(c_int) function calypso_mhd_pressure(mhd, press)        &
     &      bind(C, name="calypso_mhd_pressure")
          import :: c_int, c_double, c_ptr
          type(c_ptr), value :: mhd
          real(c_double), intent(inout) :: press(*)
        end function calypso_mhd_pressure
        integer(c_int) function calypso_mhd_energy_fluxes(mhd, nflux,    &
     &      flux_ids, fluxes) bind(C, name="calypso_mhd_energy_fluxes")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: mhd
          integer(c_int), value :: nflux
          integer(c_int), intent(in) :: flux_ids(*)
          real(c_double), intent(inout) :: fluxes(*)
        end function calypso_mhd_energy_fluxes
!
        integer(c_int) function calypso_mhd_time(mhd, time)            &
     &      bind(C, name="calypso_mhd_time")
          import :: c_int, c_double, c_ptr
          type(c_ptr), value :: mhd
          real(c_double), intent(inout) :: time
        end function calypso_mhd_time
!
        integer(c_int) function calypso_scalar_shell_create(ltr, nr,     &
     &      radius, dt, nscalar, scalars, val_icb, val_cmb, stream, sh)   &
     &      bind(C, name="calypso_scalar_shell_create")
          import :: c_int, c_double, c_ptr, calypso_scalar_bc
          integer(c_int), value :: ltr, nr, nscalar
          real(c_double), intent(in) :: radius(*), val_icb(*), val_cmb(*)
          real(c_double), value :: dt
          type(calypso_scalar_bc), intent(in) :: scalars(*)
          type(c_ptr), value :: stream
          type(c_ptr), intent(out) :: sh
        end function calypso_scalar_shell_create
!
        integer(c_int) function calypso_scalar_shell_destroy(sh)         &
     &      bind(C, name="calypso_scalar_shell_destroy")
          import :: c_int, c_ptr
          type(c_ptr), value :: sh
        end function calypso_scalar_shell_destroy
!
        integer(c_int) function calypso_scalar_shell_set_noise(sh)       &
     &      bind(C, name="calypso_scalar_shell_set_noise")
          import :: c_int, c_ptr
          type(c_ptr), value :: sh
        end function calypso_scalar_shell_set_noise
!
        integer(c_int) function calypso_scalar_shell_set_field(sh, iq,   &
     &      host) bind(C, name="calypso_scalar_shell_set_field")
          import :: c_int, c_double, c_ptr
          type(c_ptr), value :: sh
          integer(c_int), value :: iq
          real(c_double), intent(in) :: host(*)
        end function calypso_scalar_shell_set_field
!
        integer(c_int) function calypso_scalar_shell_get_field(sh, iq,   &
     &      host) bind(C, name="calypso_scalar_shell_get_field")
          import :: c_int, c_double, c_ptr
          type(c_ptr), value :: sh
          integer(c_int), value :: iq
          real(c_double), intent(inout) :: host(*)
        end function calypso_scalar_shell_get_field
!
        integer(c_int) function calypso_scalar_shell_initialize(sh)      &
     &      bind(C, name="calypso_scalar_shell_initialize")
          import :: c_int, c_ptr
          type(c_ptr), value :: sh
        end function calypso_scalar_shell_initialize
!
        integer(c_int) function calypso_scalar_shell_step(sh, nsteps)   &
     &      bind(C, name="calypso_scalar_shell_step")
          import :: c_int, c_ptr
          type(c_ptr), value :: sh
          integer(c_int), value :: nsteps
        end function calypso_scalar_shell_step
!
        integer(c_int) function calypso_scalar_shell_mean_square(sh,    &
     &      ms) bind(C, name="calypso_scalar_shell_mean_square")
          import :: c_int, c_double, c_ptr
          type(c_ptr), value :: sh
          real(c_double), intent(inout) :: ms(*)
        end function calypso_scalar_shell_mean_square
!
        integer(c_int) function calypso_scalar_shell_write_volume_ms(sh, &
     &      path) bind(C, name="calypso_scalar_shell_write_volume_ms")
          import :: c_int, c_char, c_ptr
          type(c_ptr), value :: sh
          character(kind=c_char), intent(in) :: path(*)
        end function calypso_scalar_shell_write_volume_ms
!
        integer(c_int) function calypso_mhd_write_volume_ms(mhd, path)  &
     &      bind(C, name="calypso_mhd_write_volume_ms")
          import :: c_int, c_char, c_ptr
          type(c_ptr), value :: mhd
          character(kind=c_char), intent(in) :: path(*)
        end function calypso_mhd_write_volume_ms
!
        integer(c_int) function calypso_mhd_write_restart(mhd, prefix,  &
     &      istep_fld, name_out, name_len)                              &
     &      bind(C, name="calypso_mhd_write_restart")
          import :: c_int, c_char, c_ptr
          type(c_ptr), value :: mhd
          character(kind=c_char), intent(in) :: prefix(*)
          integer(c_int), value :: istep_fld, name_len
          character(kind=c_char), intent(inout) :: name_out(*)
        end function calypso_mhd_write_restart
!
        integer(c_int) function calypso_mhd_snapshot_rtp(mhd, nvec,     &
     &      vec_comp, nscl, scl_comp, host_rtp)                         &
     &      bind(C, name="calypso_mhd_snapshot_rtp")
          import :: c_int, c_double, c_ptr
          type(c_ptr), value :: mhd
          integer(c_int), value :: nvec, nscl
          integer(c_int), intent(in) :: vec_comp(*), scl_comp(*)
          real(c_double), intent(inout) :: host_rtp(*)
        end function calypso_mhd_snapshot_rtp
!
        integer(c_int) function calypso_mhd_read_restart(mhd, path)     &
     &      bind(C, name="calypso_mhd_read_restart")
          import :: c_int, c_char, c_ptr
          type(c_ptr), value :: mhd
          character(kind=c_char), intent(in) :: path(*)
        end function calypso_mhd_read_restart
!
      end interface
!
      end module calypso_amd_capi
