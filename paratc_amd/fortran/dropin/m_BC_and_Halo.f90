! m_BC_and_Halo.f90 (drop-in) -- the reference's m_BC_and_Halo interface
! (src/m_BC_and_Halo.f90:13-17) on the device.  x and z are periodic in the
! device layout, so the halo refreshes are implicit and a download fills the
! host array's x/z halo images; these routines set the wall ghost rows.
module m_BC_and_Halo
  use iso_c_binding
  use m_paratc_gpu
  use m_paratc_resident
  implicit none
  private
  public :: SetBC_and_UpdateHalo, SetBC_and_UpdateHaloForPrSrc
  public :: SetBC_and_UpdateHalo_pr, SetBC_and_UpdateHalo_scalar
contains

  ! src/m_BC_and_Halo.f90:23-64
  subroutine SetBC_and_UpdateHalo(ux, uy, uz, uy_ym)
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: ux, uy, uz
    real(c_double), dimension(:, :), intent(in), target, contiguous :: uy_ym
    integer(c_int) :: fu, fv, fw, fp
    fu = paratc_inout(ux, PTC_F_UX)
    fv = paratc_inout(uy, PTC_F_UY)
    fw = paratc_inout(uz, PTC_F_UZ)
    fp = paratc_plane_in(uy_ym, PTC_F_UY_YM)
    call gpu_check(ptc_set_bc_and_update_halo(paratc_ctx(), fu, fv, fw, fp), 'SetBC_and_UpdateHalo')
  end subroutine SetBC_and_UpdateHalo

  ! src/m_BC_and_Halo.f90:69-103
  subroutine SetBC_and_UpdateHaloForPrSrc(ux, uy, uz, uy_ym)
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: ux, uy, uz
    real(c_double), dimension(:, :), intent(in), target, contiguous :: uy_ym
    integer(c_int) :: fu, fv, fw, fp
    fu = paratc_inout(ux, PTC_F_UX)
    fv = paratc_inout(uy, PTC_F_UY)
    fw = paratc_inout(uz, PTC_F_UZ)
    fp = paratc_plane_in(uy_ym, PTC_F_UY_YM)
    call gpu_check(ptc_set_bc_and_update_halo_for_prsrc(paratc_ctx(), fu, fv, fw, fp), &
                   'SetBC_and_UpdateHaloForPrSrc')
  end subroutine SetBC_and_UpdateHaloForPrSrc

  ! src/m_BC_and_Halo.f90:108-122
  subroutine SetBC_and_UpdateHalo_pr(pressure)
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: pressure
    call gpu_check(ptc_set_bc_and_update_halo_pr(paratc_ctx(), paratc_inout(pressure, PTC_F_PR)), &
                   'SetBC_and_UpdateHalo_pr')
  end subroutine SetBC_and_UpdateHalo_pr

  ! src/m_BC_and_Halo.f90:128-178
  subroutine SetBC_and_UpdateHalo_scalar(scalar)
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: scalar
    call gpu_check(ptc_set_bc_and_update_halo_scalar(paratc_ctx(), paratc_inout(scalar, PTC_F_SCALAR)), &
                   'SetBC_and_UpdateHalo_scalar')
  end subroutine SetBC_and_UpdateHalo_scalar

end module m_BC_and_Halo
