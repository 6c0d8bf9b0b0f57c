!> ref_harness.f90 -- TEST INFRASTRUCTURE.  A bind(C) driver that links the
!! reference's own compiled Fortran routines (built from their sources under
!! the reference tree by oracle/Makefile, output in oracle/_ref/) so the tests
!! can check the C oracle against the reference arithmetic itself.
!!
!! It only sets module state and forwards calls; every computation below is
!! done by a reference routine:
!!   construct_gauss_coefs, set_gauss_colatitude  (m_gauss_points.f90:117-195)
!!   schmidt_legendres_m                          (schmidt_fix_m.f90:62-85)
!!   set_trans_legendre_rtm                       (set_legendre_matrices.f90:80-97)
!!   legendre_b_trans_vector_org / _scalar_org    (legendre_bwd_trans_org.f90:45-197)
!!   legendre_f_trans_vector_org / _scalar_org    (legendre_fwd_trans_org.f90:45-160)
!!   ludcmp_3band, ludcmp_band, lubksb_*band_mul  (direct_solver/)
!!   rfftmi, rfftmf, rfftmb                       (fftpack5.1d.f90)
!!   write_step_data, write_field_data            (m_time_data_IO.f90,
!!                                                 field_data_IO.f90)
!!   cal_fdm2_{ICB,CMB}_free_vp / _vt             (m_coef_fdm_free_ICB.f90,
!!                                                 m_coef_fdm_free_CMB.f90)
!!   cal_sph_nod_{icb,cmb}_free_*                 (set_sph_exp_free_ICB.f90,
!!                                                 set_sph_exp_free_CMB.f90)
!!   add_fix_flux_{icb,cmb}_poisson_mat           (set_sph_scalar_mat_bc.f90)
      module calypso_ref_harness
!
      use iso_c_binding
      implicit none
!
      real(kind = 8), allocatable, save :: sp_work(:), vr_work(:)
!
      contains
!
! -----------------------------------------------------------------------
!
      subroutine ref_gauss_points(n, point, weight, colat)              &
     &          bind(C, name="ref_gauss_points")
      use m_gauss_points
      integer(c_int), value :: n
      real(c_double), intent(inout) :: point(n), weight(n), colat(n)
!
      call allocate_gauss_points(n)
      call allocate_gauss_colatitude
      call construct_gauss_coefs
      call set_gauss_colatitude
      point(1:n) = w_point(1:n)
      weight(1:n) = w_coefs(1:n)
      colat(1:n) = w_colat(1:n)
      call deallocate_gauss_colatitude
      call deallocate_gauss_points
      end subroutine ref_gauss_points
!
! -----------------------------------------------------------------------
!
      subroutine ref_schmidt_legendres_m(ltr, m, theta, p, dp)          &
     &          bind(C, name="ref_schmidt_legendres_m")
      use schmidt_fix_m
      integer(c_int), value :: ltr, m
      real(c_double), value :: theta
      real(c_double), intent(inout) :: p(0:ltr), dp(0:ltr)
      real(kind = 8) :: pmn1(0:ltr), pmp1(0:ltr), df(0:ltr+2)
!
      pmn1 = 0.0d0
      pmp1 = 0.0d0
      df = 0.0d0
      call schmidt_legendres_m(ltr, m, theta, p, dp, pmn1, pmp1, df)
      end subroutine ref_schmidt_legendres_m
!
! -----------------------------------------------------------------------
!> Set the single-domain module state the org-loop transforms read.
!! jlm(3,J): (l(l+1)+m, l, m) in rlm order; lstack(0:2L+1); P, dP(nth,J);
!! g(17,J); nthreads = np_smp (radial blocks split evenly).
      subroutine ref_leg_setup(ltr, nth, nr, jlm, lstack, radius,       &
     &          colat, weight, P, dP, g, nthreads)                      &
     &          bind(C, name="ref_leg_setup")
      use m_machine_parameter
      use m_spheric_parameter
      use m_spheric_param_smp
      use m_schmidt_poly_on_rtm
      use m_work_4_sph_trans
      use set_legendre_matrices
      integer(c_int), value :: ltr, nth, nr, nthreads
      integer(c_int), intent(in) :: jlm(3,(ltr+1)**2)
      integer(c_int), intent(in) :: lstack(0:2*ltr+1)
      real(c_double), intent(in) :: radius(nr), colat(nth), weight(nth)
      real(c_double), intent(in) :: P(nth,(ltr+1)**2)
      real(c_double), intent(in) :: dP(nth,(ltr+1)**2)
      real(c_double), intent(in) :: g(17,(ltr+1)**2)
      integer :: j, m, ip
!
      l_truncation = ltr
      nidx_rlm(1) = nr
      nidx_rlm(2) = (ltr+1)**2
      nidx_rtm(1) = nr
      nidx_rtm(2) = nth
      nidx_rtm(3) = 2*ltr + 1
      nidx_rtp(1) = nr
      nidx_rtp(2) = nth
      nidx_rtp(3) = 0
      nidx_rj(1) = nr
      nidx_rj(2) = nidx_rlm(2)
      nnod_rlm = nidx_rlm(1) * nidx_rlm(2)
      nnod_rtm = nidx_rtm(1) * nidx_rtm(2) * nidx_rtm(3)
!     strides of set_global_sph_4_rlm / set_global_sph_4_rtm
      istep_rlm(1) = nidx_rlm(2)
      istep_rlm(2) = 1
      istep_rtm(1) = nidx_rtm(2)
      istep_rtm(2) = 1
      istep_rtm(3) = nidx_rtm(1) * nidx_rtm(2)
!
      if(allocated(idx_gl_1d_rlm_j)) then
        call deallocate_sph_1d_index_rlm
        call deallocate_sph_1d_index_rtm
        call deallocate_work_4_sph_trans
        deallocate(P_rtm, dPdt_rtm, g_sph_rlm, g_sph_rj)
        call deallocate_trans_schmidt_rtm
        call deallocate_gauss_colat_rtm
        call deallocate_sph_param_smp
      end if
      np_smp = max(1, nthreads)
      call allocate_sph_1d_index_rlm
      call allocate_sph_1d_index_rtm
      call allocate_work_4_sph_trans
      call allocate_schmidt_poly_rtm
      call allocate_trans_schmidt_rtm
      call allocate_gauss_colat_rtm
      call allocate_sph_param_smp
!
      do j = 1, nidx_rlm(2)
        idx_gl_1d_rlm_j(j,1:3) = jlm(1:3,j)
      end do
      do m = 1, nidx_rtm(3)
        idx_gl_1d_rtm_m(m,1) = m
        idx_gl_1d_rtm_m(m,2) = m - 1 - ltr
      end do
      lstack_rlm(0:nidx_rtm(3)) = lstack(0:nidx_rtm(3))
      do j = 1, nidx_rlm(2)
        mdx_p_rlm_rtm(j) =  jlm(3,j) + ltr + 1
        mdx_n_rlm_rtm(j) = -jlm(3,j) + ltr + 1
      end do
      radius_1d_rlm_r(1:nr) = radius(1:nr)
      a_r_1d_rlm_r(1:nr) = 1.0d0 / radius(1:nr)
      g_colat_rtm(1:nth) = colat(1:nth)
      weight_rtm(1:nth) = weight(1:nth)
      do j = 1, nth
        asin_theta_1d_rtm(j) = 1.0d0 / sin(colat(j))
      end do
      P_rtm(1:nth,1:nidx_rlm(2)) = P(1:nth,1:nidx_rlm(2))
      dPdt_rtm(1:nth,1:nidx_rlm(2)) = dP(1:nth,1:nidx_rlm(2))
      g_sph_rlm(1:nidx_rlm(2),1:17) = transpose(g(1:17,1:nidx_rlm(2)))
      call set_trans_legendre_rtm
!     hemisphere tables of the symmetric variants (set_legendre_hemispher_rtm)
      if(allocated(Ps_rtm)) call deallocate_hemi_schmidt_rtm
      call allocate_hemi_schmidt_rtm
      call set_legendre_hemispher_rtm
!
      idx_rtm_smp_stack(0,1:3) = 0
      idx_rlm_smp_stack(0,1:2) = 0
      maxidx_rlm_smp(1) = 0
      do ip = 1, np_smp
        idx_rtm_smp_stack(ip,1) = (nr * ip) / np_smp
        idx_rlm_smp_stack(ip,1) = (nr * ip) / np_smp
        maxidx_rlm_smp(1) = max(maxidx_rlm_smp(1),                      &
     &      idx_rlm_smp_stack(ip,1) - idx_rlm_smp_stack(ip-1,1))
      end do
!     set_params_sph_trans.f90:47-55: the longest degree list of an order
      maxdegree_rlm = ltr + 1
!
      if(allocated(sp_work)) deallocate(sp_work, vr_work)
      allocate(sp_work(1), vr_work(1))
      end subroutine ref_leg_setup
!
! -----------------------------------------------------------------------
!
      subroutine ref_leg_bwd(ncomp, nvector, nscalar, sp_in, vr_out)    &
     &          bind(C, name="ref_leg_bwd")
      use m_spheric_parameter
      use legendre_bwd_trans_org
      integer(c_int), value :: ncomp, nvector, nscalar
      real(c_double), intent(in) :: sp_in(ncomp*nnod_rlm)
      real(c_double), intent(inout) :: vr_out(ncomp*nnod_rtm)
!
      if(size(sp_work) .ne. ncomp*nnod_rlm) then
        deallocate(sp_work)
        allocate(sp_work(ncomp*nnod_rlm))
      end if
      sp_work(1:ncomp*nnod_rlm) = sp_in(1:ncomp*nnod_rlm)
      vr_out(1:ncomp*nnod_rtm) = 0.0d0
      if(nvector .gt. 0) then
        call legendre_b_trans_vector_org(ncomp, nvector, sp_work, vr_out)
      end if
      if(nscalar .gt. 0) then
        call legendre_b_trans_scalar_org(ncomp, nvector, nscalar,       &
     &      sp_work, vr_out)
      end if
      end subroutine ref_leg_bwd
!
! -----------------------------------------------------------------------
!
      subroutine ref_leg_fwd(ncomp, nvector, nscalar, vr_in, sp_out)    &
     &          bind(C, name="ref_leg_fwd")
      use m_spheric_parameter
      use legendre_fwd_trans_org
      integer(c_int), value :: ncomp, nvector, nscalar
      real(c_double), intent(in) :: vr_in(ncomp*nnod_rtm)
      real(c_double), intent(inout) :: sp_out(ncomp*nnod_rlm)
!
      sp_out(1:ncomp*nnod_rlm) = 0.0d0
      if(nvector .gt. 0) then
        call legendre_f_trans_vector_org(ncomp, nvector, vr_in, sp_out)
      end if
      if(nscalar .gt. 0) then
        call legendre_f_trans_scalar_org(ncomp, nvector, nscalar,       &
     &      vr_in, sp_out)
      end if
      end subroutine ref_leg_fwd
!
! -----------------------------------------------------------------------
!> The reference's SYMMETRIC_MATMUL transforms (the variant its own tests
!! select): leg_backward_trans_sym_matmul / leg_forward_trans_sym_matmul
!! (legendre_trans_sym_matmul.f90:50-108) without the MPI waits, on
!! one-domain buffers (identity receive / send tables).  The work arrays
!! are sized for the field set as sel_init_legendre_trans does
!! (legendre_transform_select.f90:228-265).
      subroutine ref_leg_sym_init(nvector, nscalar)                     &
     &          bind(C, name="ref_leg_sym_init")
      use m_legendre_work_sym_matmul
      integer(c_int), value :: nvector, nscalar
!
      if(allocated(pol_e)) call dealloc_leg_vec_sym_matmul
      if(allocated(scl_e)) call dealloc_leg_scl_sym_matmul
      call alloc_leg_vec_sym_matmul(max(nvector,1))
      call alloc_leg_scl_sym_matmul(max(nscalar,1))
      end subroutine ref_leg_sym_init
!
      subroutine ref_leg_bwd_sym(ncomp, nvector, nscalar, sp_in, vr_out) &
     &          bind(C, name="ref_leg_bwd_sym")
      use m_spheric_parameter
      use legendre_bwd_sym_matmul
      integer(c_int), value :: ncomp, nvector, nscalar
      real(c_double), intent(in) :: sp_in(ncomp*nnod_rlm)
      real(c_double), intent(inout) :: vr_out(ncomp*nnod_rtm)
      integer(kind = kint), allocatable :: irev_rlm(:), irev_rtm(:)
      integer(kind = kint) :: i, n_WR, n_WS
!
      allocate(irev_rlm(nnod_rlm), irev_rtm(nnod_rtm))
      do i = 1, nnod_rlm
        irev_rlm(i) = i
      end do
      do i = 1, nnod_rtm
        irev_rtm(i) = i
      end do
      n_WR = ncomp*nnod_rlm
      n_WS = ncomp*nnod_rtm
      if(size(sp_work) .ne. n_WR) then
        deallocate(sp_work)
        allocate(sp_work(n_WR))
      end if
      sp_work(1:n_WR) = sp_in(1:n_WR)
      vr_out(1:n_WS) = 0.0d0
      if(nvector .gt. 0) then
        call leg_b_trans_vec_sym_matmul(ncomp, nvector,                 &
     &      irev_rlm, irev_rtm, n_WR, n_WS, sp_work, vr_out)
      end if
      if(nscalar .gt. 0) then
        call leg_b_trans_scl_sym_matmul(ncomp, nvector, nscalar,        &
     &      irev_rlm, irev_rtm, n_WR, n_WS, sp_work, vr_out)
      end if
      deallocate(irev_rlm, irev_rtm)
      end subroutine ref_leg_bwd_sym
!
      subroutine ref_leg_fwd_sym(ncomp, nvector, nscalar, vr_in, sp_out) &
     &          bind(C, name="ref_leg_fwd_sym")
      use m_spheric_parameter
      use legendre_fwd_sym_matmul
      integer(c_int), value :: ncomp, nvector, nscalar
      real(c_double), intent(in) :: vr_in(ncomp*nnod_rtm)
      real(c_double), intent(inout) :: sp_out(ncomp*nnod_rlm)
      integer(kind = kint), allocatable :: irev_rlm(:), irev_rtm(:)
      real(kind = kreal), allocatable :: wr(:)
      integer(kind = kint) :: i, n_WR, n_WS
!
      allocate(irev_rlm(nnod_rlm), irev_rtm(nnod_rtm))
      do i = 1, nnod_rlm
        irev_rlm(i) = i
      end do
      do i = 1, nnod_rtm
        irev_rtm(i) = i
      end do
      n_WR = ncomp*nnod_rtm
      n_WS = ncomp*nnod_rlm
      allocate(wr(n_WR))
      wr(1:n_WR) = vr_in(1:n_WR)
      sp_out(1:n_WS) = 0.0d0
      if(nvector .gt. 0) then
        call leg_f_trans_vec_sym_matmul(ncomp, nvector,                 &
     &      irev_rtm, irev_rlm, n_WR, n_WS, wr, sp_out)
      end if
      if(nscalar .gt. 0) then
        call leg_f_trans_scl_sym_matmul(ncomp, nvector, nscalar,        &
     &      irev_rtm, irev_rlm, n_WR, n_WS, wr, sp_out)
      end if
      deallocate(irev_rlm, irev_rtm, wr)
      end subroutine ref_leg_fwd_sym
!
! -----------------------------------------------------------------------
!> ludcmp_3band (m_ludcmp_3band.f90:76-173) on each of mcomp matrices:
!! band_a(3,n,mcomp) -> band_lu(5,n,mcomp), i_pivot(n,mcomp), ierr(mcomp)
      subroutine ref_ludcmp_3band(n, mcomp, band_a, band_lu, i_pivot,   &
     &          ierr) bind(C, name="ref_ludcmp_3band")
      use m_ludcmp_3band
      integer(c_int), value :: n, mcomp
      real(c_double), intent(in) :: band_a(3,n,mcomp)
      real(c_double), intent(inout) :: band_lu(5,n,mcomp)
      integer(c_int), intent(inout) :: i_pivot(n,mcomp), ierr(mcomp)
      real(kind = 8) :: d
      integer :: j
!
      do j = 1, mcomp
        call ludcmp_3band(n, band_a(1,1,j), i_pivot(1,j), ierr(j),      &
     &      band_lu(1,1,j), d)
      end do
      end subroutine ref_ludcmp_3band
!
! -----------------------------------------------------------------------
!> ludcmp_band (m_ludcmp_band.f90:83-207) with nband = 5:
!! band_a(5,n,mcomp) -> band_lu(9,n,mcomp), indx(n,mcomp)
      subroutine ref_ludcmp_5band(n, mcomp, band_a, band_lu, indx)      &
     &          bind(C, name="ref_ludcmp_5band")
      use m_ludcmp_band
      integer(c_int), value :: n, mcomp
      real(c_double), intent(in) :: band_a(5,n,mcomp)
      real(c_double), intent(inout) :: band_lu(9,n,mcomp)
      integer(c_int), intent(inout) :: indx(n,mcomp)
      real(kind = 8) :: d
      integer :: j
!
      do j = 1, mcomp
        call ludcmp_band(n, 5, band_a(1,1,j), band_lu(1,1,j),           &
     &      indx(1,j), d)
      end do
      end subroutine ref_ludcmp_5band
!
! -----------------------------------------------------------------------
!> lubksb_3band_mul / lubksb_5band_mul (lubksb_357band_mul.f90:73-195):
!! x(mcomp,n) solved in place, one SMP block
      subroutine ref_lubksb_mul(nb, n, mcomp, band_lu, i_pivot, x)      &
     &          bind(C, name="ref_lubksb_mul")
      use lubksb_357band_mul
      integer(c_int), value :: nb, n, mcomp
      real(c_double), intent(in) :: band_lu(*)
      integer(c_int), intent(in) :: i_pivot(n,mcomp)
      real(c_double), intent(inout) :: x(mcomp,n)
      integer(kind = 4) :: stack(0:1)
!
      stack(0) = 0
      stack(1) = mcomp
      if(nb .eq. 3) then
        call lubksb_3band_mul(1, stack, mcomp, n, band_lu, i_pivot, x)
      else
        call lubksb_5band_mul(1, stack, mcomp, n, band_lu, i_pivot, x)
      end if
      end subroutine ref_lubksb_mul
!
! -----------------------------------------------------------------------
!
! -----------------------------------------------------------------------
!
!> The reference's ASCII field file of one domain, as write_step_field_file
!! (field_file_IO.f90:40-62) writes it: write_step_data, then
!! write_field_data.  names: nfld blank-padded names of 255 characters.
      subroutine ref_write_field_file(path, rank, istep, time, dt, nnod, &
     &          nfld, ncomp, names, fdata)                                &
     &          bind(C, name="ref_write_field_file")
      use m_precision
      use m_time_data_IO
      use field_data_IO
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: rank, istep, nnod, nfld
      real(c_double), value :: time, dt
      integer(c_int), intent(in) :: ncomp(nfld)
      character(kind=c_char), intent(in) :: names(kchara, nfld)
      real(c_double), intent(in) :: fdata(nnod, *)
!
      character(len=kchara) :: fname
      character(len=kchara) :: fnames(nfld)
      integer(kind = kint) :: ncf(nfld)
      integer(kind = kint) :: i, j, ntot
!
      fname = ' '
      do i = 1, kchara
        if(path(i) .eq. c_null_char) exit
        fname(i:i) = path(i)
      end do
      do j = 1, nfld
        do i = 1, kchara
          fnames(j)(i:i) = names(i,j)
        end do
        ncf(j) = ncomp(j)
      end do
      ntot = sum(ncf)
      i_time_step_IO = istep
      time_IO = time
      delta_t_IO = dt
      open(91, file = fname, form = 'formatted')
      call write_step_data(91, rank)
      call write_field_data(91, nnod, nfld, ntot, ncf, fnames,          &
     &    fdata(1:nnod,1:ntot))
      close(91)
      end subroutine ref_write_field_file
!
! -----------------------------------------------------------------------
!> The free-slip wall tables of the four reference routines, computed as
!! set_bc_sph_mhd.f90:52-56 calls them (on radius(kr_in) and
!! radius(kr_out-1)): coef(c, pt, t), t = 1 vp ICB, 2 vt ICB, 3 vp CMB,
!! 4 vt CMB, pt -1..1 the points below / at / above the wall.
      subroutine ref_free_slip_coefs(nri, kr_in, kr_out, radius, coef)   &
     &          bind(C, name="ref_free_slip_coefs")
      use m_coef_fdm_free_ICB
      use m_coef_fdm_free_CMB
      integer(c_int), value :: nri, kr_in, kr_out
      real(c_double), intent(in) :: radius(nri)
      real(c_double), intent(inout) :: coef(3,-1:1,4)
      integer :: c, pt
!
      call cal_fdm2_ICB_free_vp(radius(kr_in))
      call cal_fdm2_ICB_free_vt(radius(kr_in))
      call cal_fdm2_CMB_free_vp(radius(kr_out-1))
      call cal_fdm2_CMB_free_vt(radius(kr_out-1))
      do pt = -1, 1
        do c = 1, 3
          coef(c,pt,1) = fdm2_free_vp_ICB(pt,c)
          coef(c,pt,2) = fdm2_free_vt_ICB(pt,c)
          coef(c,pt,3) = fdm2_free_vp_CMB(pt,c)
          coef(c,pt,4) = fdm2_free_vt_CMB(pt,c)
        end do
      end do
      end subroutine ref_free_slip_coefs
!
! -----------------------------------------------------------------------
!> One free-slip wall routine on d(nnod, 6), nnod = nri*jmax in rj order
!! (j fastest), components 1-3 the field and 4-6 its rotation or
!! diffusion; g3(jmax) = l(l+1); wall 0 the ICB (kr_in), 1 the CMB (kr_out).
!!   which 0 cal_sph_nod_*_free_v_and_w, 1 *_free_vpol2, 2 *_free_rot2,
!!         3 *_free_w_diffuse2 (coef_d; fix_fld(0:2,3,2) the three-point
!!           fixed-value tables of the ICB and the CMB)
      subroutine ref_free_slip_wall(which, wall, nri, jmax, kr_in,        &
     &          kr_out, radius, g3, coef_d, fix_fld, d)                   &
     &          bind(C, name="ref_free_slip_wall")
      use m_schmidt_poly_on_rtm
      use m_sph_spectr_data
      use m_coef_fdm_free_ICB
      use m_coef_fdm_free_CMB
      use set_sph_exp_free_ICB
      use set_sph_exp_free_CMB
      integer(c_int), value :: which, wall, nri, jmax, kr_in, kr_out
      real(c_double), value :: coef_d
      real(c_double), intent(in) :: radius(nri), g3(jmax)
      real(c_double), intent(in) :: fix_fld(0:2,3,2)
      real(c_double), intent(inout) :: d(nri*jmax,6)
      real(kind = 8) :: r_ICB(0:2), r_CMB(0:2)
!
      call cal_fdm2_ICB_free_vp(radius(kr_in))
      call cal_fdm2_ICB_free_vt(radius(kr_in))
      call cal_fdm2_CMB_free_vp(radius(kr_out-1))
      call cal_fdm2_CMB_free_vt(radius(kr_out-1))
      r_ICB(0) = radius(kr_in)
      r_ICB(1) = 1.0d0 / radius(kr_in)
      r_ICB(2) = r_ICB(1) * r_ICB(1)
      r_CMB(0) = radius(kr_out)
      r_CMB(1) = 1.0d0 / radius(kr_out)
      r_CMB(2) = r_CMB(1) * r_CMB(1)
      if(allocated(g_sph_rj)) deallocate(g_sph_rj)
      allocate(g_sph_rj(jmax,13))
      g_sph_rj = 0.0d0
      g_sph_rj(1:jmax,3) = g3(1:jmax)
      if(allocated(d_rj)) deallocate(d_rj)
      allocate(d_rj(nri*jmax,6))
      d_rj = d
!
      if(wall .eq. 0) then
        select case(which)
        case(0)
          call cal_sph_nod_icb_free_v_and_w(jmax, kr_in,                &
     &        fdm2_free_vp_ICB, fdm2_free_vt_ICB, 1, 4)
        case(1)
          call cal_sph_nod_icb_free_vpol2(jmax, kr_in,                  &
     &        fdm2_free_vp_ICB, 1)
        case(2)
          call cal_sph_nod_icb_free_rot2(jmax, kr_in, r_ICB,            &
     &        fdm2_free_vp_ICB, fdm2_free_vt_ICB, 1, 4)
        case(3)
          call cal_sph_nod_icb_free_w_diffuse2(jmax, kr_in, r_ICB,      &
     &        fix_fld(0,1,1), fdm2_free_vt_ICB, coef_d, 1, 4)
        end select
      else
        select case(which)
        case(0)
          call cal_sph_nod_cmb_free_v_and_w(jmax, kr_out,               &
     &        fdm2_free_vp_CMB, fdm2_free_vt_CMB, 1, 4)
        case(1)
          call cal_sph_nod_cmb_free_vpol2(jmax, kr_out,                 &
     &        fdm2_free_vp_CMB, 1)
        case(2)
          call cal_sph_nod_cmb_free_rot2(jmax, kr_out, r_CMB,           &
     &        fdm2_free_vp_CMB, fdm2_free_vt_CMB, 1, 4)
        case(3)
          call cal_sph_nod_cmb_free_w_diffuse2(jmax, kr_out, r_CMB,     &
     &        fix_fld(0,1,2), fdm2_free_vt_CMB, coef_d, 1, 4)
        end select
      end if
      d = d_rj
      deallocate(d_rj, g_sph_rj)
      end subroutine ref_free_slip_wall
!
! -----------------------------------------------------------------------
!> add_fix_flux_icb_poisson_mat / add_fix_flux_cmb_poisson_mat
!! (set_sph_scalar_mat_bc.f90:78-175) on mat(3, nri, jmax): wall 0 the ICB
!! row kr, 1 the CMB row kr; r3 = (r, 1/r, 1/r^2) of the wall; fdm(3,-1:1)
!! the wall table as fdm(c, pt).
      subroutine ref_add_fix_flux_rows(wall, nri, jmax, kr, r3, fdm, g3,  &
     &          coef, mat) bind(C, name="ref_add_fix_flux_rows")
      use m_schmidt_poly_on_rtm
      use set_sph_scalar_mat_bc
      integer(c_int), value :: wall, nri, jmax, kr
      real(c_double), value :: coef
      real(c_double), intent(in) :: r3(0:2), fdm(3,-1:1), g3(jmax)
      real(c_double), intent(inout) :: mat(3,nri,jmax)
      real(kind = 8) :: tab(-1:1,3)
      integer :: c, pt
!
      do pt = -1, 1
        do c = 1, 3
          tab(pt,c) = fdm(c,pt)
        end do
      end do
      if(allocated(g_sph_rj)) deallocate(g_sph_rj)
      allocate(g_sph_rj(jmax,13))
      g_sph_rj = 0.0d0
      g_sph_rj(1:jmax,3) = g3(1:jmax)
      if(wall .eq. 0) then
        call add_fix_flux_icb_poisson_mat(nri, jmax, kr, r3, tab,       &
     &      coef, mat)
      else
        call add_fix_flux_cmb_poisson_mat(nri, jmax, kr, r3, tab,       &
     &      coef, mat)
      end if
      deallocate(g_sph_rj)
      end subroutine ref_add_fix_flux_rows
!
      end module calypso_ref_harness
