!> leg_driver.f90 -- Fortran program that exercises the legendre_trans_hip
!! module end to end: it reads a grid description and an rlm spectrum from a
!! stream file, runs the GPU backward transform and then the forward
!! transform of the result, and writes both outputs.
!!
!! usage: leg_driver <mode string> <in file> <out file>
!! in file (stream, native endian): int32 L, nth, nr, nvector, nscalar,
!!   real64 radius(nr), colat(nth), weight(nth), sp(ncomp*J*nr)
!! out file: real64 vr(ncomp*nth*nr*(2L+1)), sp2(ncomp*J*nr)
      program leg_driver
!
      use iso_c_binding
      use legendre_trans_hip
      implicit none
!
      character(len=256) :: mode_ctl, fin, fout
      integer(c_int) :: ltr, nth, nr, nvec, nscl, ncomp, id_leg
      integer(c_int) :: n_rlm, n_rtm
      real(c_double), allocatable :: radius(:), colat(:), weight(:)
      real(c_double), allocatable :: sp(:), vr(:), sp2(:)
!
      call get_command_argument(1, mode_ctl)
      call get_command_argument(2, fin)
      call get_command_argument(3, fout)
!
      open(10, file=trim(fin), access='stream', form='unformatted',     &
     &     status='old')
      read(10) ltr, nth, nr, nvec, nscl
      ncomp = 3*nvec + nscl
      n_rlm = ncomp * (ltr+1)**2 * nr
      n_rtm = ncomp * nth * nr * (2*ltr+1)
      allocate(radius(nr), colat(nth), weight(nth))
      allocate(sp(n_rlm), vr(n_rtm), sp2(n_rlm))
      read(10) radius, colat, weight, sp
      close(10)
!
      id_leg = -1
      call set_legendre_trans_mode_hip(mode_ctl, id_leg)
      if(.not. is_legendre_trans_hip(id_leg)) then
        write(*,*) 'unknown Legendre mode: ', trim(mode_ctl)
        error stop 2
      end if
      call init_legendre_trans_hip(id_leg, ltr, nth, nr, radius,        &
     &    colat, weight, ncomp, nvec, nscl)
      call leg_backward_trans_hip(ncomp, nvec, nscl, n_rlm, n_rtm,      &
     &    sp, vr)
      call leg_forward_trans_hip(ncomp, nvec, nscl, n_rtm, n_rlm,       &
     &    vr, sp2)
      call finalize_legendre_trans_hip
!
      open(11, file=trim(fout), access='stream', form='unformatted',    &
     &     status='replace')
      write(11) vr, sp2
      close(11)
      write(*,'(a,i0,a,i0)') 'leg_driver done: n_rtm=', n_rtm,          &
     &    ' n_rlm=', n_rlm
!
      end program leg_driver
