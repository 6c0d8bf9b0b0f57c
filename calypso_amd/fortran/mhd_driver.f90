!> mhd_driver.f90 -- Fortran program that runs the reference's
!! Dynamobench_case1 model through the sph_analyzer_mhd_hip module:
!! L = 31, 48 x 96 grid, 48 Chebyshev intervals, r_i/r_o = 0.35,
!! E = 1e-3, modified Ra = 100, Pr = 1, Pm = 5, dt = 5e-5, Crank-Nicolson
!! weight 0.6, dynamo_benchmark_1 initial fields; prints the
!! sph_pwr_volume_s.dat row every i_step_check steps.
!!
!! usage: mhd_driver [nsteps [i_step_check]]
      program mhd_driver
!
      use iso_c_binding
      use calypso_amd_capi
      use sph_analyzer_mhd_hip
      implicit none
!
      type(calypso_mhd_params) :: p
      integer(c_int), parameter :: nri = 48
      real(c_double) :: radius(nri+1), ene(7), pi, r_icb, r_cmb, shell
      real(c_double) :: dummy(1,3), dtemp(1)
      integer(c_int) :: k, i_step, nsteps, ncheck, iflag_finish
      character(len=32) :: arg
!
      nsteps = 100
      ncheck = 50
      if(command_argument_count() .ge. 1) then
        call get_command_argument(1, arg)
        read(arg,*) nsteps
      end if
      if(command_argument_count() .ge. 2) then
        call get_command_argument(2, arg)
        read(arg,*) ncheck
      end if
!
      pi = 4.0d0 * atan(1.0d0)
      r_cmb = 1.0d0 / (1.0d0 - 0.35d0)
      r_icb = r_cmb * 0.35d0
      shell = r_cmb - r_icb
      do k = 0, nri
        radius(k+1) = r_icb + 0.5d0 * shell                             &
     &               * (1.0d0 - cos(pi * dble(k) / dble(nri)))
      end do
!
      p%l_truncation = 31
      p%nth = 48
      p%nphi = 96
      p%nr = nri + 1
      p%legendre_mode = CALYPSO_LEG_FP64
      p%dt = 5.0d-5
      p%coef_imp_v = 0.6d0
      p%coef_imp_t = 0.6d0
      p%coef_imp_b = 0.6d0
      p%coef_velo = 1.0d0
      p%coef_d_velo = 1.0d0
      p%coef_buo = 1.0d0 * (1.0d0 - 0.35d0) * 1.0d2 * 1.0d-3**(-1)
      p%coef_cor = 1.0d0 * 2.0d0 * 1.0d-3**(-1)
      p%coef_lor = 1.0d0 * 5.0d0**(-1) * 1.0d-3**(-1)
      p%coef_temp = 1.0d0
      p%coef_d_temp = 1.0d0
      p%coef_magne = 1.0d0
      p%coef_d_magne = 1.0d0 * 5.0d0**(-1)
      p%coef_induct = 1.0d0
      p%depth_high_t = 0.5384615384615384d0
      p%depth_low_t = 1.5384615384615384d0
      p%high_temp = 1.0d0
      p%low_temp = 0.0d0
      p%bc_velo_icb = CALYPSO_VELO_RIGID
      p%bc_velo_cmb = CALYPSO_VELO_RIGID
      p%kr_icb = 0
      p%kr_cmb = 0
      p%bc_magne_icb = CALYPSO_MAGNE_INSULATOR
      p%coef_press = 1.0d0 * 1.0d-3**(-1)
!
      dummy = 0.0d0
      dtemp = 0.0d0
      call SPH_initialize_MHD_hip(p, radius, 400_c_int, 1_c_int,        &
     &    dummy, dummy, dtemp)
      iflag_finish = 0
      do i_step = 1, nsteps
        call SPH_analyze_MHD_hip(i_step, nsteps, iflag_finish)
        if(mod(i_step, ncheck) .eq. 0) then
          call sph_mhd_energies_hip(ene)
          write(*,'(i16,1p8E25.15e3)') i_step, dble(i_step)*p%dt, ene
        end if
        if(iflag_finish .gt. 0) exit
      end do
      call SPH_finalize_MHD_hip
!
      end program mhd_driver
