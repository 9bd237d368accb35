! m_TScheme.f90 (drop-in) -- the reference's m_TScheme interface
! (src/m_TScheme.f90:14-19) on the device: the right-hand sides, the
! Crank-Nicolson wall-normal solves, the Poisson source and the projection
! updates.  The substep coefficients (pmGamma, pmTheta, pmAlpha, pmAlphaC,
! pmBeta, pmBetaSc, dt) are read from the host's m_Parameters at every call,
! as the reference's routines read them; PrGradAver, which clcRhsXYZ
! accumulates under IsUxConst (:284,298), is returned to m_Parameters.
! Build with -DScalarFlow for the scalar variant, as the reference.
module m_TScheme
  use iso_c_binding
  use m_paratc_gpu
  use m_paratc_resident
  use m_Parameters, only: IsUxConst
  implicit none
  private
  public :: InitTimeScheme, FluidVelUpdate, clcRhsXYZ
  public :: clcPrSrc, clcU1Hat, clcU2Hat, clcU3Hat, PressureUpdate
#ifdef ScalarFlow
  public :: clcScalar
#endif
contains

  ! InitTimeScheme (src/m_TScheme.f90:25-43)
  subroutine InitTimeScheme()
    call gpu_check(ptc_init_time_scheme(paratc_ctx()), 'InitTimeScheme')
  end subroutine InitTimeScheme

  ! clcPrSrc (src/m_TScheme.f90:48-77): prsrc = div(u)/(pmAlpha), divmax = max|div|
  subroutine clcPrSrc(ux, uy, uz, prsrc, divmax)
    real(c_double), dimension(:, :, :), intent(in), target, contiguous :: ux, uy, uz
    real(c_double), dimension(:, :, :), intent(out), target, contiguous :: prsrc
    real(c_double), intent(out) :: divmax
    integer(c_int) :: fu, fv, fw, fs
    fu = paratc_in(ux, PTC_F_UX)
    fv = paratc_in(uy, PTC_F_UY)
    fw = paratc_in(uz, PTC_F_UZ)
    fs = paratc_out(prsrc, PTC_F_ARR1)
    call paratc_push_coe()
    call gpu_check(ptc_clc_prsrc(paratc_ctx(), fu, fv, fw, fs, divmax), 'clcPrSrc')
  end subroutine clcPrSrc

  ! FluidVelUpdate (src/m_TScheme.f90:82-107)
  subroutine FluidVelUpdate(prphiHalo, ux, uy, uz)
    real(c_double), dimension(:, :, :), intent(in), target, contiguous :: prphiHalo
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: ux, uy, uz
    integer(c_int) :: ff, fu, fv, fw
    ff = paratc_in(prphiHalo, PTC_F_HALO)
    fu = paratc_inout(ux, PTC_F_UX)
    fv = paratc_inout(uy, PTC_F_UY)
    fw = paratc_inout(uz, PTC_F_UZ)
    call paratc_push_coe()
    call gpu_check(ptc_fluid_vel_update(paratc_ctx(), ff, fu, fv, fw), 'FluidVelUpdate')
  end subroutine FluidVelUpdate

#ifdef ScalarFlow
  ! clcRhsXYZ, ScalarFlow build (src/m_TScheme.f90:113-121, 130-360)
  subroutine clcRhsXYZ(ux, uy, uz, scalar, RhsX, RhsY, RhsZ, RhsC, HistXOld, HistYOld, HistZOld, HistCOld, pressure)
    real(c_double), dimension(:, :, :), intent(in), target, contiguous :: ux, uy, uz, scalar, pressure
    real(c_double), dimension(:, :, :), intent(out), target, contiguous :: RhsX, RhsY, RhsZ
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: RhsC
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: HistXOld, HistYOld, HistZOld, HistCOld
    integer(c_int) :: fu, fv, fw, fc, fp, rx, ry, rz, rc, hx, hy, hz, hc
    fu = paratc_in(ux, PTC_F_UX)
    fv = paratc_in(uy, PTC_F_UY)
    fw = paratc_in(uz, PTC_F_UZ)
    fc = paratc_in(scalar, PTC_F_SCALAR)
    fp = paratc_in(pressure, PTC_F_PR)
    rx = paratc_out(RhsX, PTC_F_ARR1)
    ry = paratc_out(RhsY, PTC_F_ARR2)
    rz = paratc_out(RhsZ, PTC_F_HALO)
    rc = paratc_inout(RhsC, PTC_F_ARRC)
    hx = paratc_inout(HistXOld, PTC_F_HISTX)
    hy = paratc_inout(HistYOld, PTC_F_HISTY)
    hz = paratc_inout(HistZOld, PTC_F_HISTZ)
    hc = paratc_inout(HistCOld, PTC_F_HISTC)
    call paratc_push_coe()
    call gpu_check(ptc_clc_rhs_xyz(paratc_ctx(), fu, fv, fw, fc, rx, ry, rz, rc, hx, hy, hz, hc, fp), 'clcRhsXYZ')
    if (IsUxConst) call paratc_pull_prgrad()
  end subroutine clcRhsXYZ
#else
  ! clcRhsXYZ (src/m_TScheme.f90:123-360)
  subroutine clcRhsXYZ(ux, uy, uz, RhsX, RhsY, RhsZ, HistXOld, HistYOld, HistZOld, pressure)
    real(c_double), dimension(:, :, :), intent(in), target, contiguous :: ux, uy, uz, pressure
    real(c_double), dimension(:, :, :), intent(out), target, contiguous :: RhsX, RhsY, RhsZ
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: HistXOld, HistYOld, HistZOld
    integer(c_int) :: fu, fv, fw, fp, rx, ry, rz, hx, hy, hz
    fu = paratc_in(ux, PTC_F_UX)
    fv = paratc_in(uy, PTC_F_UY)
    fw = paratc_in(uz, PTC_F_UZ)
    fp = paratc_in(pressure, PTC_F_PR)
    rx = paratc_out(RhsX, PTC_F_ARR1)
    ry = paratc_out(RhsY, PTC_F_ARR2)
    rz = paratc_out(RhsZ, PTC_F_HALO)
    hx = paratc_inout(HistXOld, PTC_F_HISTX)
    hy = paratc_inout(HistYOld, PTC_F_HISTY)
    hz = paratc_inout(HistZOld, PTC_F_HISTZ)
    call paratc_push_coe()
    call gpu_check(ptc_clc_rhs_xyz(paratc_ctx(), fu, fv, fw, -1_c_int, rx, ry, rz, -1_c_int, hx, hy, hz, &
                                   -1_c_int, fp), 'clcRhsXYZ')
    if (IsUxConst) call paratc_pull_prgrad()
  end subroutine clcRhsXYZ
#endif

  ! clcU1Hat (src/m_TScheme.f90:365-394): (I - pmBeta L_y) du = RhsX, ux += du
  subroutine clcU1Hat(ux, RhsX)
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: RhsX
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: ux
    integer(c_int) :: fu, fr
    fr = paratc_inout(RhsX, PTC_F_ARR1)
    fu = paratc_inout(ux, PTC_F_UX)
    call paratc_push_coe()
    call gpu_check(ptc_clc_u1hat(paratc_ctx(), fu, fr), 'clcU1Hat')
  end subroutine clcU1Hat

  ! clcU2Hat (src/m_TScheme.f90:399-437)
  subroutine clcU2Hat(uy, RhsY, duy_ym)
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: RhsY
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: uy
    real(c_double), dimension(:, :), intent(in), target, contiguous :: duy_ym
    integer(c_int) :: fv, fr, fd
    fr = paratc_inout(RhsY, PTC_F_ARR2)
    fv = paratc_inout(uy, PTC_F_UY)
    fd = paratc_plane_in(duy_ym, PTC_F_DUY_YM)
    call paratc_push_coe()
    call gpu_check(ptc_clc_u2hat(paratc_ctx(), fv, fr, fd), 'clcU2Hat')
  end subroutine clcU2Hat

  ! clcU3Hat (src/m_TScheme.f90:442-471)
  subroutine clcU3Hat(uz, RhsZ)
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: RhsZ
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: uz
    integer(c_int) :: fw, fr
    fr = paratc_inout(RhsZ, PTC_F_HALO)
    fw = paratc_inout(uz, PTC_F_UZ)
    call paratc_push_coe()
    call gpu_check(ptc_clc_u3hat(paratc_ctx(), fw, fr), 'clcU3Hat')
  end subroutine clcU3Hat

#ifdef ScalarFlow
  ! clcScalar (src/m_TScheme.f90:477-506)
  subroutine clcScalar(scalar, RhsC)
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: RhsC
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: scalar
    integer(c_int) :: fc, fr
    fr = paratc_inout(RhsC, PTC_F_ARRC)
    fc = paratc_inout(scalar, PTC_F_SCALAR)
    call paratc_push_coe()
    call gpu_check(ptc_clc_scalar(paratc_ctx(), fc, fr), 'clcScalar')
  end subroutine clcScalar
#endif

  ! PressureUpdate (src/m_TScheme.f90:512-532)
  subroutine PressureUpdate(pressure, prphiHalo)
    real(c_double), dimension(:, :, :), intent(in), target, contiguous :: prphiHalo
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: pressure
    integer(c_int) :: fp, ff
    ff = paratc_in(prphiHalo, PTC_F_HALO)
    fp = paratc_inout(pressure, PTC_F_PR)
    call paratc_push_coe()
    call gpu_check(ptc_pressure_update(paratc_ctx(), fp, ff), 'PressureUpdate')
  end subroutine PressureUpdate

end module m_TScheme
