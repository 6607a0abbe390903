! Compile-and-call check of ngw_dropin.F90 outside NanoGW (see
! ngw_dropin_selftest_types.f90): reads a small problem from a stream file
! (argument 1), calls the drop-in matvec_isdf_BLOCK and
! lanczos_spectra_isdf_BLOCK as Zcalculate_sigma_lanczos_BLOCK does, then the
! single-vector matvec_isdf, polynomial_matvec_isdf and lanczos_spectra_isdf
! as Zcalculate_sigma_lanczos does (tmp_Cmtrx built from Psi, here over the
! pair list in reverse order), then their UltraLowMem variants, then the
! lowcomsym routines on a two-irrep blocking, and writes the results to a
! stream file (argument 2).  Between the BLOCK calls and the single-vector
! ones the conduction energies in wfn%e1 are raised by 0.1 Ry in place for
! one more matvec_isdf_BLOCK call (the update an SCGW iteration makes), then
! restored: the drop-in must rebuild its resident operator both times.  After
! the lowcomsym calls one value of PsiC_intp_bl is changed in place for one
! more matvec_isdf_lowcomsym call.
program ngw_dropin_check
  use ngw_dropin_selftest_types
  use ngw_session, only: ngw_release
  implicit none
  type(ISDF) :: isdf_in
  type(wavefunction) :: wfn
  integer :: nmu, nv, nc, nb, ncv, blksz, nz, niter, npoly, icv, u
  integer, allocatable :: cvpair(:, :), tmp_cvpair(:)
  real(dp), allocatable :: vec(:, :), hvec(:, :), sqrtR(:), poly(:), v0norm(:), dummy(:)
  real(dp), allocatable :: cmtrx(:, :), sqrtr1(:), v1(:), h1(:), p1(:), veig(:)
  complex(dp), allocatable :: zz(:, :), spectra(:, :), spec1(:), speig(:)
  integer :: iv, ic, jcv
  integer, allocatable :: cvpair2(:, :)
  type(gspace) :: gvec
  integer :: nv1, nc1, nsub, jr, k
  real(dp), allocatable :: sqrtrs(:), vs(:, :), hs(:, :), hst(:, :), ps(:, :), hs2(:, :), vsn(:)
  complex(dp), allocatable :: specs(:, :), spect(:, :), zzs(:, :)
  real(dp), allocatable :: hu(:), pu(:), veig2(:)
  complex(dp), allocatable :: specu(:), speigu(:)
  character(len=512) :: fin, fout
  real(dp), allocatable :: hvec_e(:, :), sqrtr_e(:), e1_save(:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read (u) nmu, nv, nc, blksz, nz, niter, npoly
  nb = nv + nc
  ncv = nv * nc
  allocate (isdf_in%Psi_intp_loc(nmu, nb, 1, 1), isdf_in%Mmtrx_loc(nmu, nmu, 1, 1, 1, 1, 1), wfn%e1(nb))
  allocate (vec(ncv, blksz), hvec(ncv, blksz), sqrtR(ncv), poly(npoly + 1), v0norm(blksz))
  allocate (zz(nz, blksz), spectra(nz, blksz), cvpair(4, ncv), tmp_cvpair(2 * ncv), dummy(1))
  allocate (isdf_in%Mmtrx(nmu, nmu, 1, 1, 1, 1, 1), cmtrx(nmu, ncv), sqrtr1(ncv), v1(ncv), h1(ncv), p1(ncv))
  allocate (veig(ncv), spec1(nz), speig(nz))
  read (u) isdf_in%Psi_intp_loc, isdf_in%Mmtrx_loc, wfn%e1, vec, poly, zz, veig
  close (u)
  isdf_in%n_intp_r = nmu
  isdf_in%Mmtrx = isdf_in%Mmtrx_loc
  ! NanoGW's pair list: conduction outer, valence inner (setup_g.F90z:154-168)
  do icv = 1, ncv
    cvpair(1, icv) = mod(icv - 1, nv) + 1
    cvpair(2, icv) = nv + (icv - 1) / nv + 1
    cvpair(3:4, icv) = 1
    sqrtR(icv) = sqrt(wfn%e1(cvpair(2, icv)) - wfn%e1(cvpair(1, icv)))
  end do
  call matvec_isdf_BLOCK(vec, hvec, ncv, isdf_in, wfn, cvpair, ncv, sqrtR, dummy, dummy, dummy, dummy, &
                         tmp_cvpair, blksz)
  v0norm = sqrt(sum(vec**2, dim=1))
  call lanczos_spectra_isdf_BLOCK(vec, v0norm, ncv, zz, nz, niter, isdf_in, wfn, cvpair, ncv, sqrtR, poly, &
                                  npoly, spectra, dummy, dummy, dummy, dummy, tmp_cvpair, blksz)
  ! energies changed in place between calls
  allocate (hvec_e(ncv, blksz), sqrtr_e(ncv), e1_save(nb))
  e1_save = wfn%e1
  wfn%e1(nv + 1:nb) = wfn%e1(nv + 1:nb) + 0.1_dp
  do icv = 1, ncv
    sqrtr_e(icv) = sqrt(wfn%e1(cvpair(2, icv)) - wfn%e1(cvpair(1, icv)))
  end do
  call matvec_isdf_BLOCK(vec, hvec_e, ncv, isdf_in, wfn, cvpair, ncv, sqrtr_e, dummy, dummy, dummy, dummy, &
                         tmp_cvpair, blksz)
  wfn%e1 = e1_save
  ! single-vector routines: pair jcv = ncv + 1 - icv
  do icv = 1, ncv
    jcv = ncv + 1 - icv
    iv = cvpair(1, icv)
    ic = cvpair(2, icv)
    sqrtr1(jcv) = sqrt(wfn%e1(ic) - wfn%e1(iv))
    cmtrx(:, jcv) = isdf_in%Psi_intp_loc(:, ic, 1, 1) * isdf_in%Psi_intp_loc(:, iv, 1, 1)
    v1(jcv) = vec(icv, 1)
  end do
  call matvec_isdf(v1, h1, ncv, isdf_in, wfn, cmtrx, sqrtr1)
  call polynomial_matvec_isdf(v1, p1, ncv, poly, npoly, isdf_in, wfn, cmtrx, sqrtr1)
  call lanczos_spectra_isdf(v1, sqrt(sum(v1**2)), ncv, zz(:, 1), nz, niter, isdf_in, wfn, cmtrx, sqrtr1, &
                            poly, npoly, spec1)
  call lanczos_spectra_isdf(veig, sqrt(sum(veig**2)), ncv, zz(:, 1), nz, niter, isdf_in, wfn, cmtrx, sqrtr1, &
                            poly, npoly, speig)
  ! UltraLowMem variants: two-row pair list in NanoGW's order
  allocate (cvpair2(2, ncv), hu(ncv), pu(ncv), veig2(ncv), specu(nz), speigu(nz))
  cvpair2 = cvpair(1:2, :)
  do icv = 1, ncv
    veig2(icv) = veig(ncv + 1 - icv)
  end do
  call matvec_isdf_UltraLowMem(vec(:, 1), hu, ncv, isdf_in, wfn, cvpair2, sqrtR, dummy, dummy, dummy, dummy, &
                               tmp_cvpair)
  call polynomial_matvec_isdf_UltraLowMem(vec(:, 1), pu, ncv, poly, npoly, isdf_in, wfn, cvpair2, sqrtR, dummy, &
                                          dummy, dummy, dummy, tmp_cvpair)
  call lanczos_spectra_isdf_UltraLowMem(vec(:, 1), v0norm(1), ncv, zz(:, 1), nz, niter, isdf_in, wfn, cvpair2, &
                                        sqrtR, poly, npoly, specu, dummy, dummy, dummy, dummy, tmp_cvpair)
  call lanczos_spectra_isdf_UltraLowMem(veig2, sqrt(sum(veig2**2)), ncv, zz(:, 1), nz, niter, isdf_in, wfn, &
                                        cvpair2, sqrtR, poly, npoly, speigu, dummy, dummy, dummy, dummy, tmp_cvpair)
  ! lowcomsym: irreps 1 and 2 with the Z2 product table, nmrep = 1 (pairs
  ! within one irrep); valence 1..nv/2 in irrep 1, the rest in irrep 2;
  ! conduction states alternate, odd in irrep 1
  nv1 = nv / 2
  nc1 = (nc + 1) / 2
  allocate (isdf_in%PsiV_intp_bl(nmu, nv, 1, 1), isdf_in%PsiC_intp_bl(nmu, nc, 1, 1))
  isdf_in%PsiV_intp_bl(:, :, 1, 1) = isdf_in%Psi_intp_loc(:, 1:nv, 1, 1)
  isdf_in%PsiC_intp_bl(:, :, 1, 1) = isdf_in%Psi_intp_loc(:, nv + 1:nb, 1, 1)
  allocate (isdf_in%lvrep(nv), isdf_in%lcrep(nc), isdf_in%lvrep_bound(2, 2), isdf_in%lcrep_bound(2, 2))
  isdf_in%lvrep = [(iv, iv = 1, nv)]
  isdf_in%lvrep_bound(1, :) = [1, nv1]
  isdf_in%lvrep_bound(2, :) = [nv1 + 1, nv]
  isdf_in%lcrep = [(2 * ic - 1, ic = 1, nc1), (2 * ic, ic = 1, nc / 2)]
  isdf_in%lcrep_bound(1, :) = [1, nc1]
  isdf_in%lcrep_bound(2, :) = [nc1 + 1, nc]
  allocate (gvec%syms%prod(2, 2))
  gvec%syms%ntrans = 2
  gvec%syms%prod = reshape([1, 2, 2, 1], [2, 2])
  nsub = nv1 * nc1 + (nv - nv1) * (nc / 2)
  allocate (sqrtrs(nsub), vs(nsub, 2), hs(nsub, 2), hst(nsub, 2), ps(nsub, 2), vsn(2))
  allocate (specs(nz, 2), spect(nz, 2), zzs(nz, 2))
  icv = 0
  do jr = 1, 2
    do iv = isdf_in%lvrep_bound(jr, 1), isdf_in%lvrep_bound(jr, 2)
      do ic = isdf_in%lcrep_bound(jr, 1), isdf_in%lcrep_bound(jr, 2)
        icv = icv + 1
        sqrtrs(icv) = sqrt(wfn%e1(nv + isdf_in%lcrep(ic)) - wfn%e1(isdf_in%lvrep(iv)))
      end do
    end do
  end do
  do k = 1, 2
    vs(:, k) = vec(1:nsub, k)
    zzs(:, k) = zz(:, k)
  end do
  vsn = sqrt(sum(vs**2, dim=1))
  call matvec_isdf_lowcomsym(vs, hs, nsub, isdf_in, sqrtrs, .false., 2, 1, gvec)
  call matvec_isdf_lowcomsym(vs, hst, nsub, isdf_in, sqrtrs, .true., 2, 1, gvec)
  call polynomial_matvec_isdf_lowcomsym(vs, ps, nsub, poly, npoly, isdf_in, sqrtrs, 2, 1, gvec)
  call lanczos_spectra_isdf_lowcomsym(vs, vsn, nsub, zzs, nz, niter, isdf_in, sqrtrs, poly, npoly, specs, .false., &
                                      2, 1, gvec)
  call lanczos_spectra_isdf_lowcomsym(vs, vsn, nsub, zzs, nz, niter, isdf_in, sqrtrs, poly, npoly, spect, .true., &
                                      2, 1, gvec)
  ! one orbital value changed in place: the drop-in sees it and rebuilds
  allocate (hs2(nsub, 2))
  isdf_in%PsiC_intp_bl(3, 2, 1, 1) = isdf_in%PsiC_intp_bl(3, 2, 1, 1) + 0.5_dp
  call matvec_isdf_lowcomsym(vs, hs2, nsub, isdf_in, sqrtrs, .false., 2, 1, gvec)
  call ngw_release()
  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) hvec, spectra, h1, p1, spec1, speig, hu, pu, specu, speigu, hs, hst, ps, specs, spect, hs2, hvec_e
  close (u)
end program ngw_dropin_check
