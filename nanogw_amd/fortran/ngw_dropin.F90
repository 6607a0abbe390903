! Drop-in replacements of NanoGW's Lanczos hot-path routines on libngw.
!
! Link this file into NanoGW instead of shared/apply_matvec_BLOCK.F90,
! shared/lanczos_spectra_BLOCK.F90, shared/apply_matvec.f90,
! shared/lanczos_spectra.f90, shared/apply_matvec_UltraLowMem.F90,
! shared/lanczos_spectra_UltraLowMem.F90, shared/apply_matvec_lowcomsym.F90
! and shared/lanczos_spectra_lowcomsym.F90: the external subroutines keep the
! reference's names and argument lists (implicit interfaces, arguments by
! reference), so the Zcalculate_sigma_lanczos drivers (BLOCK, plain,
! UltraLowMem, lowcomsym) call them unchanged.
!
! BLOCK routines: the operator (Psi at the ISDF points, Mmtrx, energies) is
! uploaded to the GPU once and kept resident; it is re-uploaded when the
! caller passes a different array or size.  Scope: one MPI rank per GPU
! holding the whole transition space (w_grp%npes = 1), all nval x ncond pairs
! in NanoGW's conduction-outer / valence-inner order
! (shared/setup_g.F90z:154-168); other layouts stop with a message, as
! NanoGW's die() would.
!
! UltraLowMem routines (matvec_isdf_UltraLowMem and its polynomial and
! Lanczos companions): single vectors on the BLOCK routines' resident
! operator, same scope.
!
! lowcomsym routines (CPU-build argument lists): the irrep-blocked operator
! over the pairs of representation nmrep, uploaded as its orbital blocks and
! applied in factored form in the reference's pair order (see
! ngw_lowcomsym_operator).
!
! Single-vector routines (matvec_isdf, polynomial_matvec_isdf,
! lanczos_spectra_isdf): the caller's tmpCmtrx, Mmtrx and sqrtR are uploaded
! as an explicit operator (any pair list) and kept resident while the same
! arrays, sizes, sqrtR and a sample of Cmtrx and Mmtrx are passed again.
! One MPI rank (peinf%npes = 1); FP64 whatever precision the BLOCK routines
! use.
!
! Built inside the NanoGW tree it uses NanoGW's typedefs and mpi_module; with
! -DNGW_DROPIN_SELFTEST it compiles against ngw_dropin_selftest_types, a
! minimal module with the same component names, for a compile and call check
! of this file outside NanoGW.
#ifdef NGW_DROPIN_SELFTEST
#define NGW_TYPES use ngw_dropin_selftest_types
#else
#define NGW_TYPES use typedefs; use mpi_module
#endif

module ngw_session

  use iso_c_binding
  use ngw_c
  implicit none

  type(c_ptr), save :: ngw_ctx = c_null_ptr
  type(c_ptr), save :: up_psi = c_null_ptr, up_m = c_null_ptr
  integer, save :: up_nintp = -1, up_nband = -1, up_nval = -1
  ! values the resident factored operator was built from: e1(1:nband) and
  ! samples of Psi_intp and Mmtrx (in-place updates between calls, e.g. the
  ! energies an SCGW iteration rewrites, trigger a re-upload)
  real(c_double), allocatable, save :: up_e1(:), up_psamp(:), up_pmsamp(:)
  integer, save :: ngw_device = 0
  integer, save :: ngw_precision = 0
  ! explicit-operator state of the single-vector routines
  integer, parameter :: ngw_nsample = 2048
  type(c_ptr), save :: up_c = c_null_ptr, up_mx = c_null_ptr
  integer, save :: up_cnintp = -1, up_cncv = -1
  real(c_double), allocatable, save :: up_sqrtr(:), up_csamp(:), up_msamp(:)
  ! irrep-blocked (lowcomsym) operator: full copies of what was uploaded
  ! (the gathered Psi blocks, the block sizes, Mmtrx_loc of nmrep, sqrtR)
  integer, save :: up_lncv = -1
  integer(c_int), allocatable, save :: up_lnvb(:), up_lncb(:)
  real(c_double), allocatable, save :: up_lsqrtr(:), up_lpv(:, :), up_lpc(:, :), up_lmm(:, :)

contains

  subroutine ngw_start()
    if (.not. c_associated(ngw_ctx)) then
      call ngw_check(ngw_create(ngw_ctx, int(ngw_device, c_int)), 'ngw_create')
      call ngw_check(ngw_set_precision(ngw_ctx, int(ngw_precision, c_int)), 'ngw_set_precision')
    end if
  end subroutine ngw_start

  ! Evenly spaced entries of a(1:n).
  subroutine ngw_sample(a, n, out)
    integer, intent(in) :: n
    real(c_double), intent(in) :: a(n)
    real(c_double), allocatable, intent(out) :: out(:)
    integer :: m, i, stride
    m = min(n, ngw_nsample)
    stride = max(1, n / max(m, 1))
    allocate (out(m))
    do i = 1, m
      out(i) = a(1 + (i - 1) * stride)
    end do
  end subroutine ngw_sample

  ! Upload tmpCmtrx(n_intp, ncv), Mmtrx(ld_m, n_intp) and sqrtR(ncv) as the
  ! explicit operator unless the same ones are resident.
  subroutine ngw_ensure_explicit(n_intp, ncv, cmtrx, ld_m, mmtrx, sqrtr)
    integer, intent(in) :: n_intp, ncv, ld_m
    real(c_double), intent(in), target :: cmtrx(n_intp, ncv), mmtrx(ld_m, n_intp)
    real(c_double), intent(in) :: sqrtr(ncv)
    real(c_double), allocatable :: cs(:), ms(:)
    call ngw_start()
    call ngw_sample(cmtrx, n_intp * ncv, cs)
    call ngw_sample(mmtrx, ld_m * n_intp, ms)
    if (c_associated(up_c, c_loc(cmtrx(1, 1))) .and. c_associated(up_mx, c_loc(mmtrx(1, 1))) .and. &
        up_cnintp == n_intp .and. up_cncv == ncv) then
      if (all(up_sqrtr == sqrtr) .and. all(up_csamp == cs) .and. all(up_msamp == ms)) return
    end if
    call ngw_check(ngw_set_explicit_c(ngw_ctx, int(n_intp, c_int), int(ncv, c_int), cmtrx, int(n_intp, c_int), &
                                      mmtrx, int(ld_m, c_int), sqrtr), 'ngw_set_explicit_c')
    up_c = c_loc(cmtrx(1, 1))
    up_mx = c_loc(mmtrx(1, 1))
    up_cnintp = n_intp
    up_cncv = ncv
    up_sqrtr = sqrtr
    call move_alloc(cs, up_csamp)
    call move_alloc(ms, up_msamp)
    ! the factored operator is gone from the context
    up_psi = c_null_ptr
    up_m = c_null_ptr
    up_nintp = -1
    up_lncv = -1
  end subroutine ngw_ensure_explicit

  ! nval and nband from the pair list; checks NanoGW's all-pairs ordering.
  subroutine ngw_pairs(cvpair, ldc, ncv, nval, nband)
    integer, intent(in) :: ldc, ncv
    integer, intent(in) :: cvpair(ldc, ncv)
    integer, intent(out) :: nval, nband
    integer :: icv, iv, ic
    nval = maxval(cvpair(1, 1:ncv))
    nband = maxval(cvpair(2, 1:ncv))
    if (nval * (nband - nval) /= ncv) then
      write (6, *) 'libngw: pair list is not all nval x ncond pairs', nval, nband, ncv
      stop 1
    end if
    do icv = 1, ncv
      iv = mod(icv - 1, nval) + 1
      ic = nval + (icv - 1) / nval + 1
      if (cvpair(1, icv) /= iv .or. cvpair(2, icv) /= ic) then
        write (6, *) 'libngw: pair order differs from conduction-outer/valence-inner at', icv
        stop 1
      end if
    end do
  end subroutine ngw_pairs

  ! Upload the factored operator unless the same one is resident: same
  ! arrays and sizes, and the same e1(1:nband) and samples of Psi_intp and
  ! Mmtrx as at the last upload.  The library forms sqrtR from e1 as the
  ! reference drivers do (calculate_sigma_lanczos_BLOCK.F90z:199-205); a
  ! caller's sqrtR that disagrees with wfn%e1 stops with a message.
  subroutine ngw_ensure(n_intp, ld_psi, psi, ld_m, mmtrx, nband, nval, e1, ncv, sqrtr)
    integer, intent(in) :: n_intp, ld_psi, ld_m, nband, nval, ncv
    real(c_double), intent(in), target :: psi(ld_psi, *), mmtrx(ld_m, *)
    real(c_double), intent(in) :: e1(*), sqrtr(ncv)
    real(c_double), allocatable :: ps(:), ms(:)
    integer :: icv, iv, ic
    real(c_double) :: ref
    call ngw_start()
    do icv = 1, ncv
      iv = mod(icv - 1, nval) + 1
      ic = nval + (icv - 1) / nval + 1
      ref = sqrt(e1(ic) - e1(iv))
      if (abs(sqrtr(icv) - ref) > 1.0d-12 * max(1.0d0, ref)) then
        write (6, *) 'libngw: sqrtR does not match wfn%e1 at pair', icv, sqrtr(icv), ref
        stop 1
      end if
    end do
    call ngw_sample(psi, ld_psi * nband, ps)
    call ngw_sample(mmtrx, ld_m * n_intp, ms)
    if (c_associated(up_psi, c_loc(psi(1, 1))) .and. c_associated(up_m, c_loc(mmtrx(1, 1))) .and. &
        up_nintp == n_intp .and. up_nband == nband .and. up_nval == nval) then
      if (size(up_psamp) == size(ps) .and. size(up_pmsamp) == size(ms)) then
        if (all(up_e1 == e1(1:nband)) .and. all(up_psamp == ps) .and. all(up_pmsamp == ms)) return
      end if
    end if
    call ngw_check(ngw_set_isdf(ngw_ctx, int(n_intp, c_int), int(nband, c_int), int(nval, c_int), &
                                psi, int(ld_psi, c_int), mmtrx, int(ld_m, c_int), e1), 'ngw_set_isdf')
    call ngw_check(ngw_set_precision(ngw_ctx, int(ngw_precision, c_int)), 'ngw_set_precision')
    up_psi = c_loc(psi(1, 1))
    up_m = c_loc(mmtrx(1, 1))
    up_nintp = n_intp
    up_nband = nband
    up_nval = nval
    up_e1 = e1(1:nband)
    call move_alloc(ps, up_psamp)
    call move_alloc(ms, up_pmsamp)
    up_c = c_null_ptr
    up_cnintp = -1
    up_lncv = -1
  end subroutine ngw_ensure

  ! The single-vector reference Lanczos routines leave the recurrence at the
  ! first |b_k| < 1e-8 and then read a(k..niter), b(k+1..niter) unset in
  ! the continued fraction.  Given the coefficients of the full recurrence,
  ! this ends the fraction at level k - 1 instead, its exact-arithmetic value;
  ! spectra are left as they are when no b_k is that small.
  subroutine ngw_breakdown_fraction(a_elmt, b_elmt, niter, zz, nz, spectra)
    integer, intent(in) :: niter, nz
    real(c_double), intent(in) :: a_elmt(niter), b_elmt(niter)
    complex(c_double_complex), intent(in) :: zz(nz)
    complex(c_double_complex), intent(inout) :: spectra(nz)
    real(c_double), parameter :: eps = 1.0d-8
    complex(c_double_complex) :: s, enum
    integer :: k, j, iz, levels
    levels = niter
    do k = 2, niter
      if (abs(b_elmt(k)) < eps) then
        levels = k - 1
        exit
      end if
    end do
    if (levels == niter) return
    do iz = 1, nz
      s = (0.d0, 0.d0)
      do j = levels, 1, -1
        enum = b_elmt(j) * b_elmt(j)
        if (j == 1) enum = (1.d0, 0.d0)
        s = enum / (zz(iz) - a_elmt(j) - s)
      end do
      spectra(iz) = s
    end do
  end subroutine ngw_breakdown_fraction

  ! Single-vector Lanczos on the resident operator with the reference's
  ! breakdown exit applied to the fraction.
  subroutine ngw_lanczos_single(v0, v0_norm, ncv, zz, nz, niter, polynomial, npoly, spectra, what)
    integer, intent(in) :: ncv, nz, niter, npoly
    real(c_double), intent(in) :: v0(ncv), v0_norm, polynomial(npoly + 1)
    complex(c_double_complex), intent(in) :: zz(nz)
    complex(c_double_complex), intent(out) :: spectra(nz)
    character(len=*), intent(in) :: what
    real(c_double), allocatable, target :: a_elmt(:), b_elmt(:)
    real(c_double) :: vn(1)
    allocate (a_elmt(niter), b_elmt(niter))
    vn(1) = v0_norm
    call ngw_check(ngw_lanczos_spectra_isdf_block(ngw_ctx, v0, vn, int(ncv, c_int), zz, int(nz, c_int), &
                                                  int(niter, c_int), polynomial, int(npoly, c_int), spectra, &
                                                  1_c_int, 0_c_int, c_loc(a_elmt), c_loc(b_elmt)), what)
    call ngw_breakdown_fraction(a_elmt, b_elmt, niter, zz, nz, spectra)
    deallocate (a_elmt, b_elmt)
  end subroutine ngw_lanczos_single

  ! The irrep-blocked operator of matvec_isdf_lowcomsym in factored form:
  ! pv / pc hold the blocks' valence / conduction orbitals at the ISDF
  ! points, nvb / ncb the block sizes.  Uploaded unless the resident copy is
  ! equal in every value (in-place changes of the caller's arrays are seen).
  subroutine ngw_ensure_lowcomsym(n_intp, nblk, nvb, ncb, pv, pc, ld_m, mmtrx, ncv, sqrtr)
    integer, intent(in) :: n_intp, nblk, ld_m, ncv
    integer(c_int), intent(in) :: nvb(nblk), ncb(nblk)
    real(c_double), intent(in) :: pv(:, :), pc(:, :), mmtrx(ld_m, n_intp), sqrtr(ncv)
    call ngw_start()
    if (up_lncv == ncv .and. allocated(up_lnvb)) then
      if (size(up_lnvb) == nblk .and. size(up_lpv, 1) == n_intp .and. size(up_lpv, 2) == size(pv, 2) .and. &
          size(up_lpc, 2) == size(pc, 2)) then
        if (all(up_lnvb == nvb) .and. all(up_lncb == ncb) .and. all(up_lsqrtr == sqrtr) .and. &
            all(up_lpv == pv) .and. all(up_lpc == pc) .and. all(up_lmm == mmtrx(1:n_intp, 1:n_intp))) return
      end if
    end if
    call ngw_check(ngw_set_isdf_blocks(ngw_ctx, int(n_intp, c_int), int(nblk, c_int), nvb, ncb, pv, &
                                       int(n_intp, c_int), pc, int(n_intp, c_int), mmtrx, int(ld_m, c_int), sqrtr), &
                   'ngw_set_isdf_blocks (lowcomsym)')
    up_lncv = ncv
    up_lnvb = nvb
    up_lncb = ncb
    up_lsqrtr = sqrtr
    up_lpv = pv
    up_lpc = pc
    up_lmm = mmtrx(1:n_intp, 1:n_intp)
    ! the other operator forms are gone from the context
    up_psi = c_null_ptr
    up_m = c_null_ptr
    up_nintp = -1
    up_c = c_null_ptr
    up_cnintp = -1
  end subroutine ngw_ensure_lowcomsym

  subroutine ngw_release()
    integer(c_int) :: st
    if (c_associated(ngw_ctx)) st = ngw_destroy(ngw_ctx)
    ngw_ctx = c_null_ptr
    up_psi = c_null_ptr
    up_m = c_null_ptr
    up_nintp = -1
    up_c = c_null_ptr
    up_mx = c_null_ptr
    up_cnintp = -1
    up_lncv = -1
  end subroutine ngw_release

end module ngw_session

! shared/apply_matvec_BLOCK.F90:25
subroutine matvec_isdf_BLOCK(vec, Hvec, ncv_loc, isdf_in, wfn, cvpair, ncv, sqrtR, &
                             MC_vec, CMC_vec, tmp_vec, tmp_Hvec, tmp_cvpair, blksz)
  NGW_TYPES
  use ngw_session
  implicit none
  integer, intent(in) :: ncv_loc, blksz, ncv
  real(dp), intent(in) :: vec(ncv_loc, blksz), sqrtR(ncv_loc)
  real(dp), intent(out) :: Hvec(ncv_loc, blksz)
  type(ISDF), intent(in) :: isdf_in
  type(wavefunction), intent(in) :: wfn
  integer, intent(in) :: cvpair(4, ncv)
  real(dp), intent(inout) :: MC_vec(*), CMC_vec(*), tmp_vec(*), tmp_Hvec(*)
  integer, intent(inout) :: tmp_cvpair(*)
  integer :: nval, nband
  if (w_grp%npes /= 1) stop 'libngw drop-in: one rank per GPU (w_grp%npes = 1)'
  call ngw_pairs(cvpair, 4, ncv, nval, nband)
  call ngw_ensure(isdf_in%n_intp_r, size(isdf_in%Psi_intp_loc, 1), isdf_in%Psi_intp_loc(:, :, 1, 1), &
                  size(isdf_in%Mmtrx_loc, 1), isdf_in%Mmtrx_loc(:, :, 1, 1, 1, 1, 1), nband, nval, wfn%e1, &
                  ncv_loc, sqrtR)
  call ngw_check(ngw_matvec_isdf_block(ngw_ctx, vec, Hvec, int(ncv_loc, c_int), int(blksz, c_int), 0_c_int), &
                 'matvec_isdf_BLOCK')
end subroutine matvec_isdf_BLOCK

! shared/apply_matvec_BLOCK.F90:210
subroutine polynomial_matvec_isdf_BLOCK(vec, Hvec, ncv_loc, polynomial, npoly, isdf_in, wfn, cvpair, ncv, &
                                        sqrtR, MC_vec, CMC_vec, tmp_vec, tmp_Hvec, tmp_cvpair, blksz)
  NGW_TYPES
  use ngw_session
  implicit none
  integer, intent(in) :: ncv_loc, npoly, blksz, ncv
  real(dp), intent(in) :: vec(ncv_loc, blksz), polynomial(npoly + 1), sqrtR(ncv_loc)
  real(dp), intent(out) :: Hvec(ncv_loc, blksz)
  type(ISDF), intent(in) :: isdf_in
  type(wavefunction), intent(in) :: wfn
  integer, intent(in) :: cvpair(4, ncv)
  real(dp), intent(inout) :: MC_vec(*), CMC_vec(*), tmp_vec(*), tmp_Hvec(*)
  integer, intent(inout) :: tmp_cvpair(*)
  integer :: nval, nband
  if (w_grp%npes /= 1) stop 'libngw drop-in: one rank per GPU (w_grp%npes = 1)'
  call ngw_pairs(cvpair, 4, ncv, nval, nband)
  call ngw_ensure(isdf_in%n_intp_r, size(isdf_in%Psi_intp_loc, 1), isdf_in%Psi_intp_loc(:, :, 1, 1), &
                  size(isdf_in%Mmtrx_loc, 1), isdf_in%Mmtrx_loc(:, :, 1, 1, 1, 1, 1), nband, nval, wfn%e1, &
                  ncv_loc, sqrtR)
  call ngw_check(ngw_polynomial_matvec_isdf_block(ngw_ctx, vec, Hvec, int(ncv_loc, c_int), polynomial, &
                                                  int(npoly, c_int), int(blksz, c_int), 0_c_int), &
                 'polynomial_matvec_isdf_BLOCK')
end subroutine polynomial_matvec_isdf_BLOCK

! shared/lanczos_spectra_BLOCK.F90:7
subroutine lanczos_spectra_isdf_BLOCK(v0, v0_norm, ncv_loc, zz, nz, niter, isdf_in, wfn, cvpair, ncv, sqrtR, &
                                      polynomial, npoly, spectra, MC_vec, CMC_vec, tmp_vec, tmp_Hvec, &
                                      tmp_cvpair, blksz)
  NGW_TYPES
  use ngw_session
  implicit none
  integer, intent(in) :: ncv_loc, niter, nz, npoly, blksz, ncv
  complex(dp), intent(in) :: zz(nz, blksz)
  real(dp), intent(in) :: v0(ncv_loc, blksz), polynomial(npoly + 1), v0_norm(blksz), sqrtR(ncv_loc)
  complex(dp), intent(out) :: spectra(nz, blksz)
  type(ISDF), intent(in) :: isdf_in
  type(wavefunction), intent(in) :: wfn
  integer, intent(in) :: cvpair(4, ncv)
  real(dp), intent(inout) :: MC_vec(*), CMC_vec(*), tmp_vec(*), tmp_Hvec(*)
  integer, intent(inout) :: tmp_cvpair(*)
  integer :: nval, nband
  if (w_grp%npes /= 1) stop 'libngw drop-in: one rank per GPU (w_grp%npes = 1)'
  call ngw_pairs(cvpair, 4, ncv, nval, nband)
  call ngw_ensure(isdf_in%n_intp_r, size(isdf_in%Psi_intp_loc, 1), isdf_in%Psi_intp_loc(:, :, 1, 1), &
                  size(isdf_in%Mmtrx_loc, 1), isdf_in%Mmtrx_loc(:, :, 1, 1, 1, 1, 1), nband, nval, wfn%e1, &
                  ncv_loc, sqrtR)
  call ngw_check(ngw_lanczos_spectra_isdf_block(ngw_ctx, v0, v0_norm, int(ncv_loc, c_int), zz, int(nz, c_int), &
                                                int(niter, c_int), polynomial, int(npoly, c_int), spectra, &
                                                int(blksz, c_int), 0_c_int, c_null_ptr, c_null_ptr), &
                 'lanczos_spectra_isdf_BLOCK')
end subroutine lanczos_spectra_isdf_BLOCK

! shared/apply_matvec.f90:25
subroutine matvec_isdf(vec, Hvec, ncv_loc, isdf_in, wfn, tmpCmtrx, sqrtR)
  NGW_TYPES
  use ngw_session
  implicit none
  integer, intent(in) :: ncv_loc
  real(dp), intent(in) :: vec(ncv_loc)
  type(ISDF), intent(in) :: isdf_in
  real(dp), intent(in) :: tmpCmtrx(isdf_in%n_intp_r, ncv_loc), sqrtR(ncv_loc)
  real(dp), intent(out) :: Hvec(ncv_loc)
  type(wavefunction), intent(in) :: wfn
  if (peinf%npes /= 1) stop 'libngw drop-in: one rank per GPU (peinf%npes = 1)'
  call ngw_ensure_explicit(isdf_in%n_intp_r, ncv_loc, tmpCmtrx, size(isdf_in%Mmtrx, 1), &
                           isdf_in%Mmtrx(:, :, 1, 1, 1, 1, 1), sqrtR)
  call ngw_check(ngw_matvec_isdf_block(ngw_ctx, vec, Hvec, int(ncv_loc, c_int), 1_c_int, 0_c_int), 'matvec_isdf')
end subroutine matvec_isdf

! shared/apply_matvec.f90:142 (tmp_Cmtrx is the (n_intp_r, ncv_loc) matrix the
! reference passes through to matvec_isdf)
subroutine polynomial_matvec_isdf(vec, Hvec, ncv_loc, polynomial, npoly, isdf_in, wfn, tmp_Cmtrx, sqrtR)
  NGW_TYPES
  use ngw_session
  implicit none
  integer, intent(in) :: ncv_loc, npoly
  type(ISDF), intent(in) :: isdf_in
  real(dp), intent(in) :: polynomial(npoly + 1), tmp_Cmtrx(isdf_in%n_intp_r, ncv_loc), sqrtR(ncv_loc)
  real(dp), intent(in) :: vec(ncv_loc)
  type(wavefunction), intent(in) :: wfn
  real(dp), intent(out) :: Hvec(ncv_loc)
  if (peinf%npes /= 1) stop 'libngw drop-in: one rank per GPU (peinf%npes = 1)'
  call ngw_ensure_explicit(isdf_in%n_intp_r, ncv_loc, tmp_Cmtrx, size(isdf_in%Mmtrx, 1), &
                           isdf_in%Mmtrx(:, :, 1, 1, 1, 1, 1), sqrtR)
  call ngw_check(ngw_polynomial_matvec_isdf_block(ngw_ctx, vec, Hvec, int(ncv_loc, c_int), polynomial, &
                                                  int(npoly, c_int), 1_c_int, 0_c_int), 'polynomial_matvec_isdf')
end subroutine polynomial_matvec_isdf

! shared/lanczos_spectra.f90:7 (breakdown exit :58, see ngw_breakdown_fraction)
subroutine lanczos_spectra_isdf(v0, v0_norm, ncv_loc, zz, nz, niter, isdf_in, &
                                wfn, tmp_Cmtrx, sqrtR, polynomial, npoly, spectra)
  NGW_TYPES
  use ngw_session
  implicit none
  integer, intent(in) :: ncv_loc, niter, nz, npoly
  complex(dp), intent(in) :: zz(nz)
  type(ISDF), intent(in) :: isdf_in
  real(dp), intent(in) :: v0(ncv_loc), polynomial(npoly + 1), v0_norm, &
                          tmp_Cmtrx(isdf_in%n_intp_r, ncv_loc), sqrtR(ncv_loc)
  complex(dp), intent(out) :: spectra(nz)
  type(wavefunction), intent(in) :: wfn
  if (peinf%npes /= 1) stop 'libngw drop-in: one rank per GPU (peinf%npes = 1)'
  call ngw_ensure_explicit(isdf_in%n_intp_r, ncv_loc, tmp_Cmtrx, size(isdf_in%Mmtrx, 1), &
                           isdf_in%Mmtrx(:, :, 1, 1, 1, 1, 1), sqrtR)
  call ngw_lanczos_single(v0, v0_norm, ncv_loc, zz, nz, niter, polynomial, npoly, spectra, 'lanczos_spectra_isdf')
end subroutine lanczos_spectra_isdf

! shared/apply_matvec_UltraLowMem.F90:25 (the pair list has two rows)
subroutine matvec_isdf_UltraLowMem(vec, Hvec, ncv_loc, isdf_in, wfn, cvpair, sqrtR, &
                                   MC_vec, CMC_vec, tmp_vec, tmp_Hvec, tmp_cvpair)
  NGW_TYPES
  use ngw_session
  implicit none
  integer, intent(in) :: ncv_loc
  real(dp), intent(in) :: vec(ncv_loc), sqrtR(ncv_loc)
  real(dp), intent(out) :: Hvec(ncv_loc)
  type(ISDF), intent(in) :: isdf_in
  type(wavefunction), intent(in) :: wfn
  integer, intent(in) :: cvpair(2, ncv_loc)
  real(dp), intent(inout) :: MC_vec(*), CMC_vec(*), tmp_vec(*), tmp_Hvec(*)
  integer, intent(inout) :: tmp_cvpair(*)
  integer :: nval, nband
  if (w_grp%npes /= 1) stop 'libngw drop-in: one rank per GPU (w_grp%npes = 1)'
  call ngw_pairs(cvpair, 2, ncv_loc, nval, nband)
  call ngw_ensure(isdf_in%n_intp_r, size(isdf_in%Psi_intp_loc, 1), isdf_in%Psi_intp_loc(:, :, 1, 1), &
                  size(isdf_in%Mmtrx_loc, 1), isdf_in%Mmtrx_loc(:, :, 1, 1, 1, 1, 1), nband, nval, wfn%e1, &
                  ncv_loc, sqrtR)
  call ngw_check(ngw_matvec_isdf_block(ngw_ctx, vec, Hvec, int(ncv_loc, c_int), 1_c_int, 0_c_int), &
                 'matvec_isdf_UltraLowMem')
end subroutine matvec_isdf_UltraLowMem

! shared/apply_matvec_UltraLowMem.F90:168
subroutine polynomial_matvec_isdf_UltraLowMem(vec, Hvec, ncv_loc, polynomial, npoly, isdf_in, wfn, cvpair, sqrtR, &
                                              MC_vec, CMC_vec, tmp_vec, tmp_Hvec, tmp_cvpair)
  NGW_TYPES
  use ngw_session
  implicit none
  integer, intent(in) :: ncv_loc, npoly
  real(dp), intent(in) :: vec(ncv_loc), polynomial(npoly + 1), sqrtR(ncv_loc)
  real(dp), intent(out) :: Hvec(ncv_loc)
  type(ISDF), intent(in) :: isdf_in
  type(wavefunction), intent(in) :: wfn
  integer, intent(in) :: cvpair(2, ncv_loc)
  real(dp), intent(inout) :: MC_vec(*), CMC_vec(*), tmp_vec(*), tmp_Hvec(*)
  integer, intent(inout) :: tmp_cvpair(*)
  integer :: nval, nband
  if (w_grp%npes /= 1) stop 'libngw drop-in: one rank per GPU (w_grp%npes = 1)'
  call ngw_pairs(cvpair, 2, ncv_loc, nval, nband)
  call ngw_ensure(isdf_in%n_intp_r, size(isdf_in%Psi_intp_loc, 1), isdf_in%Psi_intp_loc(:, :, 1, 1), &
                  size(isdf_in%Mmtrx_loc, 1), isdf_in%Mmtrx_loc(:, :, 1, 1, 1, 1, 1), nband, nval, wfn%e1, &
                  ncv_loc, sqrtR)
  call ngw_check(ngw_polynomial_matvec_isdf_block(ngw_ctx, vec, Hvec, int(ncv_loc, c_int), polynomial, &
                                                  int(npoly, c_int), 1_c_int, 0_c_int), &
                 'polynomial_matvec_isdf_UltraLowMem')
end subroutine polynomial_matvec_isdf_UltraLowMem

! shared/lanczos_spectra_UltraLowMem.F90:7 (breakdown exit :66, see
! ngw_breakdown_fraction; cvpair is (2, ncv_loc) as the driver passes it)
subroutine lanczos_spectra_isdf_UltraLowMem(v0, v0_norm, ncv_loc, zz, nz, niter, isdf_in, &
                                            wfn, cvpair, sqrtR, polynomial, npoly, spectra, &
                                            MC_vec, CMC_vec, tmp_vec, tmp_Hvec, tmp_cvpair)
  NGW_TYPES
  use ngw_session
  implicit none
  integer, intent(in) :: ncv_loc, niter, nz, npoly
  complex(dp), intent(in) :: zz(nz)
  type(ISDF), intent(in) :: isdf_in
  real(dp), intent(in) :: v0(ncv_loc), polynomial(npoly + 1), v0_norm, sqrtR(ncv_loc)
  integer, intent(in) :: cvpair(2, ncv_loc)
  complex(dp), intent(out) :: spectra(nz)
  type(wavefunction), intent(in) :: wfn
  real(dp), intent(inout) :: MC_vec(*), CMC_vec(*), tmp_vec(*), tmp_Hvec(*)
  integer, intent(inout) :: tmp_cvpair(*)
  integer :: nval, nband
  if (w_grp%npes /= 1) stop 'libngw drop-in: one rank per GPU (w_grp%npes = 1)'
  call ngw_pairs(cvpair, 2, ncv_loc, nval, nband)
  call ngw_ensure(isdf_in%n_intp_r, size(isdf_in%Psi_intp_loc, 1), isdf_in%Psi_intp_loc(:, :, 1, 1), &
                  size(isdf_in%Mmtrx_loc, 1), isdf_in%Mmtrx_loc(:, :, 1, 1, 1, 1, 1), nband, nval, wfn%e1, &
                  ncv_loc, sqrtR)
  call ngw_lanczos_single(v0, v0_norm, ncv_loc, zz, nz, niter, polynomial, npoly, spectra, &
                          'lanczos_spectra_isdf_UltraLowMem')
end subroutine lanczos_spectra_isdf_UltraLowMem

! The irrep-blocked operator of matvec_isdf_lowcomsym
! (shared/apply_matvec_lowcomsym.F90:155-412) in factored form: for each
! valence irrep jvrep the conduction irrep jcrep with prod(jcrep, jvrep) =
! nmrep is one block; its orbitals are gathered through lvrep / lcrep
! (PsiV_intp_bl(:, lvrep(iv)), PsiC_intp_bl(:, lcrep(ic))), and pairs run
! block after block, valence outer, conduction inner (the reference's
! cvstart order, :159-247).  The pair matrix is not formed.  A single rank
! (w_grp%npes = 1).
subroutine ngw_lowcomsym_operator(ncv_loc, isdf_in, sqrtR, nmrep, gvec)
  NGW_TYPES
  use ngw_session
  implicit none
  integer, intent(in) :: ncv_loc, nmrep
  type(ISDF), intent(in) :: isdf_in
  real(dp), intent(in) :: sqrtR(ncv_loc)
  type(gspace), intent(in) :: gvec
  real(dp), allocatable :: pv(:, :), pc(:, :)
  integer(c_int), allocatable :: nvb(:), ncb(:), jvs(:), jcs(:)
  integer :: n, ntr, nblk, jvrep, jcrep, iv, ic, kv, kc, k
  integer(8) :: npair
  if (w_grp%npes /= 1) stop 'libngw drop-in: one rank per GPU (w_grp%npes = 1)'
  n = isdf_in%n_intp_r
  ntr = gvec%syms%ntrans
  allocate (nvb(ntr), ncb(ntr), jvs(ntr), jcs(ntr))
  nblk = 0
  npair = 0
  do jvrep = 1, ntr
    do jcrep = 1, ntr
      if (gvec%syms%prod(jcrep, jvrep) == nmrep) exit
    end do
    if (jcrep > ntr) cycle
    nblk = nblk + 1
    jvs(nblk) = jvrep
    jcs(nblk) = jcrep
    nvb(nblk) = max(0, isdf_in%lvrep_bound(jvrep, 2) - isdf_in%lvrep_bound(jvrep, 1) + 1)
    ncb(nblk) = max(0, isdf_in%lcrep_bound(jcrep, 2) - isdf_in%lcrep_bound(jcrep, 1) + 1)
    npair = npair + int(nvb(nblk), 8) * int(ncb(nblk), 8)
  end do
  if (npair /= int(ncv_loc, 8)) stop 'libngw drop-in: lowcomsym pair count differs from ncv_loc'
  allocate (pv(n, sum(nvb(1:nblk))), pc(n, sum(ncb(1:nblk))))
  kv = 0
  kc = 0
  do k = 1, nblk
    do iv = isdf_in%lvrep_bound(jvs(k), 1), isdf_in%lvrep_bound(jvs(k), 2)
      kv = kv + 1
      pv(:, kv) = isdf_in%PsiV_intp_bl(1:n, isdf_in%lvrep(iv), 1, 1)
    end do
    do ic = isdf_in%lcrep_bound(jcs(k), 1), isdf_in%lcrep_bound(jcs(k), 2)
      kc = kc + 1
      pc(:, kc) = isdf_in%PsiC_intp_bl(1:n, isdf_in%lcrep(ic), 1, 1)
    end do
  end do
  call ngw_ensure_lowcomsym(n, nblk, nvb(1:nblk), ncb(1:nblk), pv, pc, size(isdf_in%Mmtrx_loc, 1), &
                            isdf_in%Mmtrx_loc(:, :, 1, 1, 1, 1, nmrep), ncv_loc, sqrtR)
  deallocate (pv, pc, nvb, ncb, jvs, jcs)
end subroutine ngw_lowcomsym_operator

! shared/apply_matvec_lowcomsym.F90:31 (CPU build's argument list)
subroutine matvec_isdf_lowcomsym(vec, Hvec, ncv_loc, isdf_in, sqrtR, tamm_d, blksz, nmrep, gvec)
  NGW_TYPES
  use ngw_session
  implicit none
  integer, intent(in) :: ncv_loc, blksz, nmrep
  type(gspace), intent(in) :: gvec
  real(dp), intent(in) :: vec(ncv_loc, blksz), sqrtR(ncv_loc)
  type(ISDF), intent(in) :: isdf_in
  real(dp), intent(out) :: Hvec(ncv_loc, blksz)
  logical, intent(in) :: tamm_d
  integer(c_int) :: td
  call ngw_lowcomsym_operator(ncv_loc, isdf_in, sqrtR, nmrep, gvec)
  td = 0
  if (tamm_d) td = 1
  call ngw_check(ngw_matvec_isdf_block(ngw_ctx, vec, Hvec, int(ncv_loc, c_int), int(blksz, c_int), td), &
                 'matvec_isdf_lowcomsym')
end subroutine matvec_isdf_lowcomsym

! shared/apply_matvec_lowcomsym.F90:419 (always the full operator, :478-483)
subroutine polynomial_matvec_isdf_lowcomsym(vec, Hvec, ncv_loc, polynomial, npoly, isdf_in, sqrtR, &
                                            blksz, nmrep, gvec)
  NGW_TYPES
  use ngw_session
  implicit none
  integer, intent(in) :: ncv_loc, npoly, blksz, nmrep
  type(gspace), intent(in) :: gvec
  real(dp), intent(in) :: vec(ncv_loc, blksz), polynomial(npoly + 1), sqrtR(ncv_loc)
  type(ISDF), intent(in) :: isdf_in
  real(dp), intent(out) :: Hvec(ncv_loc, blksz)
  call ngw_lowcomsym_operator(ncv_loc, isdf_in, sqrtR, nmrep, gvec)
  call ngw_check(ngw_polynomial_matvec_isdf_block(ngw_ctx, vec, Hvec, int(ncv_loc, c_int), polynomial, &
                                                  int(npoly, c_int), int(blksz, c_int), 0_c_int), &
                 'polynomial_matvec_isdf_lowcomsym')
end subroutine polynomial_matvec_isdf_lowcomsym

! shared/lanczos_spectra_lowcomsym.F90:7.  TDA runs the recurrence on the TDA
! operator itself (:76-86, :130-140), the full mode on p(H) (:87-97); no
! breakdown exit (:120).
subroutine lanczos_spectra_isdf_lowcomsym(v0, v0_norm, ncv_loc, zz, nz, niter, isdf_in, sqrtR, polynomial, npoly, &
                                          spectra, tamm_d, blksz, nmrep, gvec)
  NGW_TYPES
  use ngw_session
  implicit none
  integer, intent(in) :: ncv_loc, niter, nz, npoly, blksz, nmrep
  complex(dp), intent(in) :: zz(nz, blksz)
  type(ISDF), intent(in) :: isdf_in
  real(dp), intent(in) :: v0(ncv_loc, blksz), polynomial(npoly + 1), v0_norm(blksz), sqrtR(ncv_loc)
  type(gspace), intent(in) :: gvec
  logical, intent(in) :: tamm_d
  complex(dp), intent(out) :: spectra(nz, blksz)
  real(c_double) :: identity(2)
  call ngw_lowcomsym_operator(ncv_loc, isdf_in, sqrtR, nmrep, gvec)
  if (tamm_d) then
    identity = [1.d0, 0.d0]
    call ngw_check(ngw_lanczos_spectra_isdf_block(ngw_ctx, v0, v0_norm, int(ncv_loc, c_int), zz, int(nz, c_int), &
                                                  int(niter, c_int), identity, 1_c_int, spectra, int(blksz, c_int), &
                                                  1_c_int, c_null_ptr, c_null_ptr), 'lanczos_spectra_isdf_lowcomsym')
  else
    call ngw_check(ngw_lanczos_spectra_isdf_block(ngw_ctx, v0, v0_norm, int(ncv_loc, c_int), zz, int(nz, c_int), &
                                                  int(niter, c_int), polynomial, int(npoly, c_int), spectra, &
                                                  int(blksz, c_int), 0_c_int, c_null_ptr, c_null_ptr), &
                 'lanczos_spectra_isdf_lowcomsym')
  end if
end subroutine lanczos_spectra_isdf_lowcomsym
