! host_modules.f90 -- the host-side modules a drop-in test program needs:
! the module variables the reference's time-step routines read, under the
! names of src/m_Parameters.f90, src/m_MeshAndMetries.f90 and
! src/m_Variables.f90, which the drop-in modules (paratc_amd/fortran/dropin)
! read through `use`.  A real host uses the reference's own modules instead.
module m_Parameters
  implicit none
  integer, parameter :: FT_CH = 1, FT_HC = 2
  integer, save :: FlowType, nxc, nyc, nzc, ischeme, iCFL, iadvance
  logical, save :: IsUxConst
  real(8), save :: ubulk, xlx, yly, zlz, xnu, uCRF, dtMax, CFLc, dt, SimTime, PrGradAver
  real(8), save :: gravity(3), uxBcValue(6), uyBcValue(6), uzBcValue(6), vel_limit, div_limit
  real(8), save :: pmGamma, pmTheta, pmAlpha, pmAlphaC, pmBeta
  real(8), save :: pmGammaConst(3), pmThetaConst(3), pmAlphaConst(3)
#ifdef ScalarFlow
  real(8), save :: pmBetaSc, SDiffCoe, GravityEffVec(3), FallingVelVec(3), ScalarBcValues(4)
  integer, save :: ScalarBcOption(2)
#endif
contains

  ! the scheme tables of ReadAndInitParameters (src/m_Parameters.f90:159-174)
  subroutine set_scheme_tables()
    pmGammaConst = 0d0
    pmThetaConst = 0d0
    select case (ischeme)
    case (1)
      iadvance = 1
      pmGammaConst(1) = 1.5d0; pmThetaConst(1) = -0.5d0
    case (2)
      iadvance = 2
      pmGammaConst(1:2) = [0.5d0, 1d0]; pmThetaConst(1:2) = [0d0, -0.5d0]
    case default
      iadvance = 3
      pmGammaConst = [8d0 / 15d0, 5d0 / 12d0, 3d0 / 4d0]
      pmThetaConst = [0d0, -17d0 / 60d0, -5d0 / 12d0]
    end select
    pmAlphaConst = pmGammaConst + pmThetaConst
  end subroutine set_scheme_tables

  ! PMcoeUpdate (src/m_Parameters.f90:215-230)
  subroutine PMcoeUpdate(ns)
    integer, intent(in) :: ns
    pmGamma = pmGammaConst(ns) * dt
    pmTheta = pmThetaConst(ns) * dt
    pmAlpha = pmAlphaConst(ns) * dt
    pmBeta = 0.5d0 * pmAlphaConst(ns) * xnu * dt
#ifdef ScalarFlow
    pmBetaSc = 0.5d0 * pmAlphaConst(ns) * SDiffCoe * dt
#endif
    SimTime = SimTime + dt * pmAlphaConst(ns)
    pmAlphaC = pmAlphaConst(ns)
    if (ns == 1) PrGradAver = 0d0
  end subroutine PMcoeUpdate
end module m_Parameters

module m_MeshAndMetries
  implicit none
  real(8), allocatable, save :: yp(:)   ! yp(0:nyp), nodes in yp(1:nyp)
end module m_MeshAndMetries

! the arrays of src/m_Variables.f90:10-34 with the bounds of AllocateVariables
! on one rank (mb1: x -1..nx+2, y 0..ny+1, z -1..nz+2; y1start:y1end interior)
module m_Variables
  implicit none
  real(8), allocatable, save, dimension(:, :, :) :: ux, uy, uz, pressure, RealHalo
  real(8), allocatable, save, dimension(:, :, :) :: HistxOld, HistyOld, HistzOld, RealArr1, RealArr2
  real(8), allocatable, save, dimension(:, :) :: uy_ym, duy_ym
#ifdef ScalarFlow
  real(8), allocatable, save, dimension(:, :, :) :: scalar, HistCOld, RealArrC
#endif
contains
  subroutine AllocateVariables(nx, ny, nz)
    integer, intent(in) :: nx, ny, nz
    allocate(ux(-1:nx + 2, 0:ny + 1, -1:nz + 2), uy(-1:nx + 2, 0:ny + 1, -1:nz + 2))
    allocate(uz(-1:nx + 2, 0:ny + 1, -1:nz + 2), pressure(-1:nx + 2, 0:ny + 1, -1:nz + 2))
    allocate(RealHalo(-1:nx + 2, 0:ny + 1, -1:nz + 2))
    allocate(HistxOld(nx, ny, nz), HistyOld(nx, ny, nz), HistzOld(nx, ny, nz))
    allocate(RealArr1(nx, ny, nz), RealArr2(nx, ny, nz))
    allocate(uy_ym(nx, nz), duy_ym(nx, nz))
    ux = 0d0; uy = 0d0; uz = 0d0; pressure = 0d0; RealHalo = 0d0
    HistxOld = 0d0; HistyOld = 0d0; HistzOld = 0d0; RealArr1 = 0d0; RealArr2 = 0d0
    uy_ym = 0d0; duy_ym = 0d0
#ifdef ScalarFlow
    allocate(scalar(-1:nx + 2, 0:ny + 1, -1:nz + 2), RealArrC(-1:nx + 2, 0:ny + 1, -1:nz + 2))
    allocate(HistCOld(nx, ny, nz))
    scalar = 0d0; RealArrC = 0d0; HistCOld = 0d0
#endif
  end subroutine AllocateVariables
end module m_Variables
