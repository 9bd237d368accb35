! m_Poisson.f90 (drop-in) -- the reference's m_Poisson interface
! (src/m_Poisson.f90:22) on the device solver.  Same module and procedure
! names and arguments; the host arrays are mapped to device fields by
! m_paratc_resident.
module m_Poisson
  use iso_c_binding
  use m_paratc_gpu
  use m_paratc_resident
  implicit none
  private
  public :: InitPoissonSolver, clcPPE, Destory_Poisson_FFT_Plan
contains

  ! InitPoissonSolver (src/m_Poisson.f90:28-175): FFT plans and the timed
  ! choice of variant; leaves both arrays zero, as the reference does (:174),
  ! and prints each candidate's time and the pick, as the reference prints
  ! its choices' (:50-171)
  subroutine InitPoissonSolver(prsrc, prphiHalo)
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: prsrc, prphiHalo
    integer(c_int) :: fs, fd, n, i, i0
    character(kind=c_char) :: none(1)
    character(kind=c_char), allocatable :: buf(:)
    fs = paratc_out(prsrc, PTC_F_ARR1)
    fd = paratc_out(prphiHalo, PTC_F_HALO)
    call gpu_check(ptc_init_poisson_solver(paratc_ctx(), fs, fd), 'InitPoissonSolver')
    write(*, *) 'Auto-tuning mode for Poisson Solver......'
    call gpu_check(ptc_poisson_search_log(paratc_ctx(), none, 0_c_int, n), 'InitPoissonSolver')
    allocate(buf(n + 1))
    call gpu_check(ptc_poisson_search_log(paratc_ctx(), buf, n + 1_c_int, n), 'InitPoissonSolver')
    i0 = 1
    do i = 1, n
      if (buf(i) == c_new_line) then
        write(*, '(A,999A)') '  ', buf(i0:i - 1)
        i0 = i + 1
      end if
    end do
    deallocate(buf)
  end subroutine InitPoissonSolver

  ! clcPPE (src/m_Poisson.f90:16, the solve of :352-452): the interior of
  ! prphiHalo = L^-1 prsrc, the (kx, kz) = (0, 0) line pinned at j = 1
  subroutine clcPPE(prsrc, prphiHalo)
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: prsrc
    real(c_double), dimension(:, :, :), intent(out), target, contiguous :: prphiHalo
    integer(c_int) :: fs, fd
    fs = paratc_in(prsrc, PTC_F_ARR1)
    fd = paratc_out(prphiHalo, PTC_F_HALO)
    call gpu_check(ptc_clc_ppe(paratc_ctx(), fs, fd), 'clcPPE')
  end subroutine clcPPE

  ! Destory_Poisson_FFT_Plan (src/m_Poisson.f90:339-347)
  subroutine Destory_Poisson_FFT_Plan()
    call gpu_check(ptc_destroy_poisson_plan(paratc_ctx()), 'Destory_Poisson_FFT_Plan')
  end subroutine Destory_Poisson_FFT_Plan

end module m_Poisson
