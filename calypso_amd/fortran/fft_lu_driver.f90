!> fft_lu_driver.f90 -- Fortran program that exercises the mhd_fft_select_hip,
!! cal_nonlinear_sph_mhd_hip and lubksb_band_hip modules end to end: a
!! backward and a forward FFT through the selector, the grid products of the
!! backward result (when it has the 13 fields of the MHD step) and one band
!! solve over a range of modes.
!!
!! usage: fft_lu_driver <in file> <out file>
!! in file (stream, native endian):
!!   int32 L, nth, nr, nphi, ncomp_bwd, ncomp_fwd
!!   real64 WR(ncomp_bwd*nnod_rtm), frc_rtp(nnod_rtp*ncomp_fwd)
!!   int32 nb, n, mcomp, Msmp, Msmp_stack(0:Msmp)
!!   real64 band_lu((2nb-1)*n*mcomp), int32 i_pivot(n*mcomp),
!!   real64 x(mcomp*n)
!! out file: real64 fld_rtp(nnod_rtp*ncomp_bwd), WS(ncomp_fwd*nnod_rtm),
!!   x(mcomp*n), and if ncomp_bwd = 13 frc_rtp(nnod_rtp*19) = the products of
!!   fld_rtp with coefficients (-1.0, 0.7, 1.3, -0.4)
      program fft_lu_driver
!
      use iso_c_binding
      use mhd_fft_select_hip
      use lubksb_band_hip
      use cal_nonlinear_sph_mhd_hip
      implicit none
!
      character(len=256) :: fin, fout
      integer(c_int) :: ltr, nth, nr, nphi, ncb, ncf, nrtm, nrtp
      integer(c_int) :: nb, n, mcomp, msmp
      integer(c_int), allocatable :: stack(:), piv(:,:)
      real(c_double), allocatable :: WR(:), frc(:), fld(:), WS(:)
      real(c_double), allocatable :: lu(:,:,:), x(:,:), frc2(:)
!
      call get_command_argument(1, fin)
      call get_command_argument(2, fout)
!
      open(10, file=trim(fin), access='stream', form='unformatted',     &
     &     status='old')
      read(10) ltr, nth, nr, nphi, ncb, ncf
      nrtm = nth * nr * (2*ltr+1)
      nrtp = nth * nr * nphi
      allocate(WR(ncb*nrtm), frc(nrtp*ncf))
      allocate(fld(nrtp*ncb), WS(ncf*nrtm))
      read(10) WR, frc
      read(10) nb, n, mcomp, msmp
      allocate(stack(0:msmp))
      read(10) stack
      allocate(lu(2*nb-1,n,mcomp), piv(n,mcomp), x(mcomp,n))
      read(10) lu, piv, x
      close(10)
!
      call init_MHD_FFT_select_hip(0, ltr, nth, nr, nphi,               &
     &    max(ncb, ncf), ncf, ncb)
      call back_MHD_FFT_sel_from_recv_hip(ncb, ncb*nrtm, WR, fld)
      call fwd_MHD_FFT_sel_from_recv_hip(ncf, ncf*nrtm, frc, WS)
      call finalize_MHD_FFT_select_hip
      if(ncb .eq. 13) then
        allocate(frc2(nrtp*19))
        call s_cal_nonlinear_sph_MHD_hip(nrtp, fld, frc2,               &
     &      -1.0d0, 0.7d0, 1.3d0, -0.4d0)
      end if
!
      if(nb .eq. 3) then
        call lubksb_3band_mul_hip(msmp, stack, mcomp, n, lu, piv, x)
      else
        call lubksb_5band_mul_hip(msmp, stack, mcomp, n, lu, piv, x)
      end if
!
      open(11, file=trim(fout), access='stream', form='unformatted',    &
     &     status='replace')
      write(11) fld, WS, x
      if(ncb .eq. 13) write(11) frc2
      close(11)
      write(*,'(a,i0,a,i0)') 'fft_lu_driver done: nnod_rtp=', nrtp,     &
     &    ' mcomp=', mcomp
!
      end program fft_lu_driver
