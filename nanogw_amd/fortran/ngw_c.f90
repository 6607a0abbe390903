! iso_c_binding interfaces of libngw (include/ngw.h) for Fortran hosts.
!
! NanoGW is Fortran 90; its Lanczos drivers (SIGMA/calculate_sigma_lanczos*.F90z)
! reach the gfx950 path through these bind(C) interfaces.  All arrays are
! passed by address (assumed size), scalars by value, complex(dp) arrays as
! interleaved real pairs; ngw_check turns a non-zero status into the same
! behaviour as NanoGW's die() (shared/die.F90:10): message, then stop.
module ngw_c

  use iso_c_binding
  implicit none

  integer(c_int), parameter :: NGW_OK = 0
  integer(c_int), parameter :: NGW_PREC_F64 = 0, NGW_PREC_F32 = 1, &
                               NGW_PREC_BF16 = 2, NGW_PREC_BF16X3 = 3, NGW_PREC_F16 = 4

  interface

    function ngw_last_error() bind(C, name='ngw_last_error') result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function ngw_last_error

    integer(c_int) function ngw_create(ctx, device) bind(C, name='ngw_create')
      import :: c_ptr, c_int
      type(c_ptr), intent(out) :: ctx
      integer(c_int), value :: device
    end function ngw_create

    integer(c_int) function ngw_destroy(ctx) bind(C, name='ngw_destroy')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function ngw_destroy

    integer(c_int) function ngw_set_isdf(ctx, n_intp, nband, nval, psi_intp, ld_psi, &
                                         mmtrx, ld_m, e1) bind(C, name='ngw_set_isdf')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n_intp, nband, nval, ld_psi, ld_m
      real(c_double), intent(in) :: psi_intp(*), mmtrx(*), e1(*)
    end function ngw_set_isdf

    integer(c_int) function ngw_set_precision(ctx, precision) bind(C, name='ngw_set_precision')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: precision
    end function ngw_set_precision
    integer(c_int) function ngw_set_reorthogonalization(ctx, mode) bind(C, name='ngw_set_reorthogonalization')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
    end function ngw_set_reorthogonalization

    integer(c_int) function ngw_set_explicit_c(ctx, n_intp, ncv, cmtrx, ld_c, mmtrx, ld_m, sqrtr) &
      bind(C, name='ngw_set_explicit_c')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n_intp, ncv, ld_c, ld_m
      real(c_double), intent(in) :: cmtrx(*), mmtrx(*), sqrtr(*)
    end function ngw_set_explicit_c

    integer(c_int) function ngw_set_isdf_blocks(ctx, n_intp, nblock, nv_blk, nc_blk, psiv, ld_v, psic, ld_c, &
                                                mmtrx, ld_m, sqrtr) bind(C, name='ngw_set_isdf_blocks')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: n_intp, nblock, ld_v, ld_c, ld_m
      integer(c_int), intent(in) :: nv_blk(*), nc_blk(*)
      real(c_double), intent(in) :: psiv(*), psic(*), mmtrx(*), sqrtr(*)
    end function ngw_set_isdf_blocks

    integer(c_int) function ngw_matvec_isdf_block(ctx, vec, hvec, ncv, blksz, tamm_d) &
      bind(C, name='ngw_matvec_isdf_block')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: vec(*)
      real(c_double), intent(out) :: hvec(*)
      integer(c_int), value :: ncv, blksz, tamm_d
    end function ngw_matvec_isdf_block

    integer(c_int) function ngw_polynomial_matvec_isdf_block(ctx, vec, hvec, ncv, polynomial, &
                                                             npoly, blksz, tamm_d) &
      bind(C, name='ngw_polynomial_matvec_isdf_block')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: vec(*), polynomial(*)
      real(c_double), intent(out) :: hvec(*)
      integer(c_int), value :: ncv, npoly, blksz, tamm_d
    end function ngw_polynomial_matvec_isdf_block

    integer(c_int) function ngw_lanczos_spectra_isdf_block(ctx, v0, v0_norm, ncv, zz, nz, niter, &
                                                           polynomial, npoly, spectra, blksz, &
                                                           tamm_d, a_elmt, b_elmt) &
      bind(C, name='ngw_lanczos_spectra_isdf_block')
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: v0(*), v0_norm(*), polynomial(*)
      complex(c_double_complex), intent(in) :: zz(*)
      complex(c_double_complex), intent(out) :: spectra(*)
      integer(c_int), value :: ncv, nz, niter, npoly, blksz, tamm_d
      type(c_ptr), value :: a_elmt, b_elmt
    end function ngw_lanczos_spectra_isdf_block

    integer(c_int) function ngw_sigma_lanczos(ctx, nqp, qp_band, nen, deltae, ecuts, niter, &
                                              polynomial, npoly, tamm_d, max_cols, sigma_c) &
      bind(C, name='ngw_sigma_lanczos')
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: nqp, nen, niter, npoly, tamm_d, max_cols
      integer(c_int), intent(in) :: qp_band(*)
      real(c_double), value :: deltae, ecuts
      real(c_double), intent(in) :: polynomial(*)
      complex(c_double_complex), intent(out) :: sigma_c(*)
    end function ngw_sigma_lanczos

    integer(c_int) function ngw_sigma_lanczos_bands(ctx, nqp, qp_band, band_lo, band_hi, nen, deltae, &
                                                    ecuts, niter, polynomial, npoly, tamm_d, max_cols, &
                                                    sigma_c) bind(C, name='ngw_sigma_lanczos_bands')
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: nqp, band_lo, band_hi, nen, niter, npoly, tamm_d, max_cols
      integer(c_int), intent(in) :: qp_band(*)
      real(c_double), value :: deltae, ecuts
      real(c_double), intent(in) :: polynomial(*)
      complex(c_double_complex), intent(out) :: sigma_c(*)
    end function ngw_sigma_lanczos_bands

    integer(c_int) function ngw_sigma_x_isdf(ctx, nqp, qp_band, sigx) bind(C, name='ngw_sigma_x_isdf')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: nqp
      integer(c_int), intent(in) :: qp_band(*)
      real(c_double), intent(out) :: sigx(*)
    end function ngw_sigma_x_isdf

    integer(c_int) function ngw_qp_newton(ctx, nqp, nen, deltae_ev, edft, en0, sigx, vxc, &
                                          sigma_c, eqp, zfac, sigma0, imsig, ierr) &
      bind(C, name='ngw_qp_newton')
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: nqp, nen
      real(c_double), value :: deltae_ev
      real(c_double), intent(in) :: edft(*), en0(*), sigx(*), vxc(*)
      complex(c_double_complex), intent(in) :: sigma_c(*)
      real(c_double), intent(out) :: eqp(*), zfac(*), imsig(*)
      complex(c_double_complex), intent(out) :: sigma0(*)
      integer(c_int), intent(out) :: ierr(*)
    end function ngw_qp_newton
    ! no counterpart in NanoGW: device-resident inputs and outputs, allocates nothing
    integer(c_int) function ngw_qp_newton_dev_noalloc(ctx, nqp, nen, deltae_ev, d_in, d_out) &
      bind(C, name='ngw_qp_newton_dev_noalloc')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx, d_in, d_out
      integer(c_int), value :: nqp, nen
      real(c_double), value :: deltae_ev
    end function ngw_qp_newton_dev_noalloc

    integer(c_int) function ngw_polyfit_sqrt(xmin, xmax, npoly, nx, chebyshev, beta, fit_error) &
      bind(C, name='ngw_polyfit_sqrt')
      import :: c_int, c_double
      real(c_double), value :: xmin, xmax
      integer(c_int), value :: npoly, nx, chebyshev
      real(c_double), intent(out) :: beta(*), fit_error
    end function ngw_polyfit_sqrt

    integer(c_int) function ngw_isdf_coulomb_kernel(ctx, ngrid, n_intp, zeta, ld_zeta, grid, nfft, &
                                                    h, trunc, ntrans, mmtrx, ld_m) &
      bind(C, name='ngw_isdf_coulomb_kernel')
      import :: c_ptr, c_int, c_long, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ngrid, n_intp, ld_m
      integer(c_long), value :: ld_zeta
      real(c_double), intent(in) :: zeta(*)
      integer(c_int), intent(in) :: grid(*), nfft(*)
      real(c_double), value :: h, trunc, ntrans
      real(c_double), intent(out) :: mmtrx(*)
    end function ngw_isdf_coulomb_kernel

    integer(c_int) function ngw_isdf_zeta(ctx, ngrid, nband, nval, psi, ld_psi, n_intp, intp, ridge, &
                                          zeta, ld_zeta, used_lu) bind(C, name='ngw_isdf_zeta')
      import :: c_ptr, c_int, c_long, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ngrid, nband, nval, n_intp
      integer(c_long), value :: ld_psi, ld_zeta
      real(c_double), intent(in) :: psi(*)
      integer(c_int), intent(in) :: intp(*)
      real(c_double), value :: ridge
      real(c_double), intent(out) :: zeta(*)
      integer(c_int), intent(out) :: used_lu
    end function ngw_isdf_zeta

  end interface

contains

  ! Message of the last failed call on this thread.
  function ngw_error_text() result(msg)
    character(len=512) :: msg
    character(kind=c_char), pointer :: cs(:)
    integer :: i
    type(c_ptr) :: p
    msg = ' '
    p = ngw_last_error()
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [512])
    do i = 1, 512
      if (cs(i) == c_null_char) exit
      msg(i:i) = cs(i)
    end do
  end function ngw_error_text

  ! NanoGW's die() semantics for a failed libngw call.
  subroutine ngw_check(status, where)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: where
    if (status /= NGW_OK) then
      write (6, '(3a,i3,2a)') ' ERROR in ', where, ' (status ', status, '): ', trim(ngw_error_text())
      stop 1
    end if
  end subroutine ngw_check

end module ngw_c
