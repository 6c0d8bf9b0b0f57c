!> scalar_shell_driver.f90 -- Fortran program that runs the scalar model of
!! the reference's tests/heterogineous_temp through the bind(C) interfaces of
!! calypso_amd_capi: scalars diffusing in a shell at rest (r_i/r_o = 0.35,
!! Chebyshev radial grid), per-mode wall values or fluxes, noise initial
!! fields; appends the sph_pwr_volume_s.dat row every ncheck steps.
!!
!! usage: scalar_shell_driver <setup file> <energy file> [nsteps [ncheck]]
!! The setup file (free format): l_truncation nri dt nscalar, then per
!! scalar a line "coef_d coef_imp bc_icb bc_cmb", its (L+1)^2 inner-wall
!! values and its (L+1)^2 outer-wall values (mode l(l+1)+m), as the
!! spectral boundary file gives them.
      program scalar_shell_driver
!
      use iso_c_binding
      use calypso_amd_capi
      implicit none
!
      integer(c_int) :: ltr, nri, nsc, jmax, k, iq, i_step, nsteps, ncheck
      integer(c_int) :: ierr
      real(c_double) :: dt, pi, r_icb, r_cmb, shell
      real(c_double), allocatable :: radius(:), vi(:,:), vc(:,:)
      type(calypso_scalar_bc), allocatable :: bcs(:)
      character(len=32), allocatable, target :: names(:)
      type(c_ptr) :: hs
      character(len=512) :: setup, efile
      character(len=32) :: arg
!
      if(command_argument_count() .lt. 2) then
        write(*,*) 'usage: scalar_shell_driver <setup> <energy file> [nsteps [ncheck]]'
        error stop 2
      end if
      call get_command_argument(1, setup)
      call get_command_argument(2, efile)
      nsteps = 100
      ncheck = 50
      if(command_argument_count() .ge. 3) then
        call get_command_argument(3, arg)
        read(arg,*) nsteps
      end if
      if(command_argument_count() .ge. 4) then
        call get_command_argument(4, arg)
        read(arg,*) ncheck
      end if
!
      open(10, file=trim(setup), status='old', action='read')
      read(10,*) ltr, nri, dt, nsc
      jmax = (ltr + 1)**2
      allocate(radius(nri+1), vi(jmax,nsc), vc(jmax,nsc), bcs(nsc), names(nsc))
      do iq = 1, nsc
        read(10,*) bcs(iq)%coef_d, bcs(iq)%coef_imp, bcs(iq)%bc_icb, bcs(iq)%bc_cmb
        read(10,*) vi(:,iq)
        read(10,*) vc(:,iq)
        names(iq) = c_null_char
        bcs(iq)%name = c_null_ptr
      end do
      close(10)
      if(nsc .ge. 1) then
        names(1) = 'temperature' // c_null_char
        bcs(1)%name = c_loc(names(1))
      end if
      if(nsc .ge. 2) then
        names(2) = 'composition' // c_null_char
        bcs(2)%name = c_loc(names(2))
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
      ierr = calypso_scalar_shell_create(ltr, nri + 1, radius, dt, nsc, &
     &         bcs, vi, vc, c_null_ptr, hs)
      if(ierr .ne. 0) error stop 'calypso_scalar_shell_create failed'
      ierr = calypso_scalar_shell_set_noise(hs)
      if(ierr .ne. 0) error stop 'calypso_scalar_shell_set_noise failed'
      ierr = calypso_scalar_shell_initialize(hs)
      if(ierr .ne. 0) error stop 'calypso_scalar_shell_initialize failed'
      do i_step = 1, nsteps
        ierr = calypso_scalar_shell_step(hs, 1)
        if(ierr .ne. 0) error stop 'calypso_scalar_shell_step failed'
        if(mod(i_step, ncheck) .eq. 0) then
          ierr = calypso_scalar_shell_write_volume_ms(hs,                &
     &             trim(efile) // c_null_char)
          if(ierr .ne. 0) error stop 'calypso_scalar_shell_write_volume_ms failed'
        end if
      end do
      ierr = calypso_scalar_shell_destroy(hs)
!
      end program scalar_shell_driver
