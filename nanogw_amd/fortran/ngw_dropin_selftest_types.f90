! Stand-in types for the compile-and-call check of ngw_dropin.F90 outside NanoGW.
!
! ngw_dropin_selftest_types carries only the components of NanoGW's
! type(ISDF), type(wavefunction) and w_grp that the drop-in routines read,
! under the same names.  The program reads a small problem from a stream file
! (argument 1), calls the drop-in matvec_isdf_BLOCK and
! lanczos_spectra_isdf_BLOCK exactly as Zcalculate_sigma_lanczos_BLOCK does,
! and writes Hvec and the spectra to a stream file (argument 2).
module ngw_dropin_selftest_types
  implicit none
  integer, parameter :: dp = kind(1.0d0)
  type ISDF
    integer :: n_intp_r
    real(dp), pointer :: Psi_intp_loc(:, :, :, :)
    real(dp), pointer :: Mmtrx_loc(:, :, :, :, :, :, :)
  end type ISDF
  type wavefunction
    real(dp), pointer :: e1(:)
  end type wavefunction
  type grp_t
    integer :: npes = 1
  end type grp_t
  type(grp_t) :: w_grp
end module ngw_dropin_selftest_types

