! Stand-in types for the compile-and-call check of ngw_dropin.F90 outside NanoGW.
!
! ngw_dropin_selftest_types carries only the components of NanoGW's
! type(ISDF), type(wavefunction), type(gspace), w_grp and peinf that the drop-in routines
! read, under the same names (ngw_dropin_check.f90 is the program that uses
! them).
module ngw_dropin_selftest_types
  implicit none
  integer, parameter :: dp = kind(1.0d0)
  type ISDF
    integer :: n_intp_r
    real(dp), pointer :: Psi_intp_loc(:, :, :, :)
    real(dp), pointer :: Mmtrx_loc(:, :, :, :, :, :, :)
    real(dp), pointer :: Mmtrx(:, :, :, :, :, :, :)
    real(dp), pointer :: PsiV_intp_bl(:, :, :, :), PsiC_intp_bl(:, :, :, :)
    integer, pointer :: lvrep(:), lvrep_bound(:, :), lcrep(:), lcrep_bound(:, :)
  end type ISDF
  type symmetries
    integer, pointer :: prod(:, :)
    integer :: ntrans
  end type symmetries
  type gspace
    type(symmetries) :: syms
  end type gspace
  type wavefunction
    real(dp), pointer :: e1(:)
  end type wavefunction
  type grp_t
    integer :: npes = 1
  end type grp_t
  type(grp_t) :: w_grp, peinf
end module ngw_dropin_selftest_types

