! Compile-and-call check of ngw_dropin.F90 outside NanoGW (see
! ngw_dropin_selftest_types.f90): reads a small problem from a stream file
! (argument 1), calls the drop-in matvec_isdf_BLOCK and
! lanczos_spectra_isdf_BLOCK as Zcalculate_sigma_lanczos_BLOCK does, then the
! single-vector matvec_isdf, polynomial_matvec_isdf and lanczos_spectra_isdf
! as Zcalculate_sigma_lanczos does (tmp_Cmtrx built from Psi, here over the
! pair list in reverse order), then their UltraLowMem variants, and writes
! the results to a stream file (argument 2).
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
  real(dp), allocatable :: hu(:), pu(:), veig2(:)
  complex(dp), allocatable :: specu(:), speigu(:)
  character(len=512) :: fin, fout

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
  call ngw_release()
  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) hvec, spectra, h1, p1, spec1, speig, hu, pu, specu, speigu
  close (u)
end program ngw_dropin_check
