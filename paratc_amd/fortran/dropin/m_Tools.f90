! m_Tools.f90 (drop-in) -- the monitors of the reference's m_Tools
! (src/m_Tools.f90:24-110, 256-301) on the device, under the reference's
! names and arguments, so ChannelIterate's checks (src/m_ChannelSystem.f90:
! 93-100, 141-145) read the device fields.  The reference's other m_Tools
! routines (the vortex criteria, dissipation rate and host tridiagonal
! solvers) are not on the time-step path: a host keeps them by renaming its
! m_Tools to m_Tools_host and building this module with -DPTC_HOST_TOOLS,
! which re-exports them.
module m_Tools
  use iso_c_binding
  use m_paratc_gpu
  use m_paratc_resident
#ifdef PTC_HOST_TOOLS
  use m_Tools_host, only: Clc_Q_vor, Clc_lamda2, CalcDissipationRate, InverseTridiagonal, &
                          InversePeriodicTridiagonal, InversePTriFixedCoe
#endif
  implicit none
  private
  public :: CalcMaxCFL, CheckDivergence, CalcVmax, CalcUxAver
#ifdef ScalarFlow
  public :: ClcMeanValue
#endif
#ifdef PTC_HOST_TOOLS
  public :: Clc_Q_vor, Clc_lamda2, CalcDissipationRate, InverseTridiagonal
  public :: InversePeriodicTridiagonal, InversePTriFixedCoe
#endif
contains

  ! CalcMaxCFL (src/m_Tools.f90:24-49)
  subroutine CalcMaxCFL(ux, uy, uz, uddxmax)
    real(c_double), dimension(:, :, :), intent(in), target, contiguous :: ux, uy, uz
    real(c_double), intent(out) :: uddxmax
    integer(c_int) :: fu, fv, fw
    fu = paratc_in(ux, PTC_F_UX)
    fv = paratc_in(uy, PTC_F_UY)
    fw = paratc_in(uz, PTC_F_UZ)
    call gpu_check(ptc_calc_max_cfl(paratc_ctx(), fu, fv, fw, uddxmax), 'CalcMaxCFL')
  end subroutine CalcMaxCFL

  ! CheckDivergence (src/m_Tools.f90:54-81)
  subroutine CheckDivergence(ux, uy, uz, divmax)
    real(c_double), dimension(:, :, :), intent(in), target, contiguous :: ux, uy, uz
    real(c_double), intent(out) :: divmax
    integer(c_int) :: fu, fv, fw
    fu = paratc_in(ux, PTC_F_UX)
    fv = paratc_in(uy, PTC_F_UY)
    fw = paratc_in(uz, PTC_F_UZ)
    call gpu_check(ptc_check_divergence(paratc_ctx(), fu, fv, fw, divmax), 'CheckDivergence')
  end subroutine CheckDivergence

  ! CalcVmax (src/m_Tools.f90:86-110)
  function CalcVmax(ux, uy, uz) result(vmaxabs)
    real(c_double), dimension(:, :, :), intent(in), target, contiguous :: ux, uy, uz
    real(c_double), dimension(3) :: vmaxabs
    integer(c_int) :: fu, fv, fw
    fu = paratc_in(ux, PTC_F_UX)
    fv = paratc_in(uy, PTC_F_UY)
    fw = paratc_in(uz, PTC_F_UZ)
    call gpu_check(ptc_calc_vmax(paratc_ctx(), fu, fv, fw, vmaxabs), 'CalcVmax')
  end function CalcVmax

  ! CalcUxAver (src/m_Tools.f90:256-276)
  function CalcUxAver(ux) result(res)
    real(c_double), dimension(:, :, :), intent(in), target, contiguous :: ux
    real(c_double) :: res
    call gpu_check(ptc_calc_ux_aver(paratc_ctx(), paratc_in(ux, PTC_F_UX), res), 'CalcUxAver')
  end function CalcUxAver

#ifdef ScalarFlow
  ! ClcMeanValue (src/m_Tools.f90:281-301)
  function ClcMeanValue(ux, scalar) result(res)
    real(c_double), dimension(:, :, :), intent(in), target, contiguous :: ux, scalar
    real(c_double), dimension(3) :: res
    integer(c_int) :: fu, fc
    fu = paratc_in(ux, PTC_F_UX)
    fc = paratc_in(scalar, PTC_F_SCALAR)
    call gpu_check(ptc_clc_mean_value(paratc_ctx(), fu, fc, res), 'ClcMeanValue')
  end function ClcMeanValue
#endif

end module m_Tools
