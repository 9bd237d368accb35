! m_paratc_gpu.f90 -- Fortran binding of the HIP C-ABI (include/paratc_gpu.h).
!
! A Fortran host keeps its own main program, namelist reading, mesh set-up and
! output (src/main.f90, m_Parameters, m_MeshAndMetries, m_IOAndVisu) and calls
! the device hot path through this module.  The interfaces are one-to-one with
! the C entry points; the wrappers below take the host's arrays in the
! reference's own shapes (haloed mb1 arrays and interior y1start:y1end arrays,
! src/m_Variables.f90:47-50) and abort through the caller-supplied error path
! when a call fails, as MainLog%CheckForError(ErrT_Abort, ...) would
! (src/m_LogInfo.f90:164-170).
module m_paratc_gpu
  use iso_c_binding
  implicit none
  private

  ! field ids (paratc_gpu.h)
  integer(c_int), parameter, public :: PTC_F_UX = 0, PTC_F_UY = 1, PTC_F_UZ = 2, PTC_F_PR = 3
  integer(c_int), parameter, public :: PTC_F_SCALAR = 4, PTC_F_HISTX = 5, PTC_F_HISTY = 6
  integer(c_int), parameter, public :: PTC_F_HISTZ = 7, PTC_F_HISTC = 8, PTC_F_ARR1 = 9
  integer(c_int), parameter, public :: PTC_F_ARR2 = 10, PTC_F_HALO = 11, PTC_F_ARRC = 12
  integer(c_int), parameter, public :: PTC_F_WORK1 = 13, PTC_F_WORK2 = 14
  integer(c_int), parameter, public :: PTC_F_UY_YM = 15, PTC_F_DUY_YM = 16
  integer(c_int), parameter, public :: PTC_L_HALO = 0, PTC_L_INTERIOR = 1, PTC_L_PLANE = 2

  type, bind(C), public :: ptc_grid_t
    integer(c_int) :: nx, ny, nz, pad_
    real(c_double) :: xlx, yly, zlz
  end type ptc_grid_t

  type, bind(C), public :: ptc_params_t
    integer(c_int) :: flow_type, scalar, is_ux_const, ischeme, icfl
    integer(c_int) :: sbc_opt(2), pad_
    real(c_double) :: ubulk, xnu, ucrf, dtmax, cflc
    real(c_double) :: gravity(3)
    real(c_double) :: uxbc(6), uybc(6), uzbc(6)
    real(c_double) :: sdiff
    real(c_double) :: gev(3), fvv(3)
    real(c_double) :: sbc_val(4)
    real(c_double) :: vel_limit, div_limit
  end type ptc_params_t

  public :: ptc_create, ptc_destroy, ptc_upload, ptc_download, ptc_last_error_f
  public :: ptc_init_poisson_solver, ptc_clc_ppe, ptc_destroy_poisson_plan
  public :: ptc_set_bc_and_update_halo, ptc_set_bc_and_update_halo_pr
  public :: ptc_set_bc_and_update_halo_scalar, ptc_channel_iterate, ptc_calc_ux_aver
  public :: ptc_check_divergence, ptc_synchronize
  public :: ptc_set_coe, ptc_get_coe
  public :: ptc_pm_coe_update, ptc_init_time_scheme, ptc_clc_rhs_xyz, ptc_clc_u1hat, ptc_clc_u2hat
  public :: ptc_clc_u3hat, ptc_clc_scalar, ptc_clc_prsrc, ptc_fluid_vel_update, ptc_pressure_update
  public :: ptc_set_bc_and_update_halo_for_prsrc, ptc_update_uy_ym, ptc_calc_max_cfl, ptc_calc_vmax
  public :: ptc_clc_mean_value, ptc_poisson_config, ptc_poisson_set, ptc_poisson_get_config
  public :: ptc_poisson_search_log
  public :: ptc_poisson_get_knob, ptc_clc_stat, ptc_get_stat, ptc_visu_field
  public :: ptc_init_spectra, ptc_clc_spectra, ptc_get_spectra
  integer(c_int), parameter, public :: PTC_VISU_UX = 0, PTC_VISU_UY = 1, PTC_VISU_UZ = 2, PTC_VISU_PR = 3
  integer(c_int), parameter, public :: PTC_VISU_WX = 4, PTC_VISU_WY = 5, PTC_VISU_WZ = 6, PTC_VISU_WMAG = 7
  integer(c_int), parameter, public :: PTC_VISU_SCALAR = 8, PTC_VISU_Q = 9, PTC_VISU_LAMBDA2 = 10
  public :: gpu_check, gpu_put, gpu_get, gpu_clcPPE

  integer(c_int), parameter, public :: PTC_PREC_FP64 = 0, PTC_PREC_FP32_REFINE = 1
  integer(c_int), parameter, public :: PTC_TRI_AUTO = 0, PTC_TRI_THOMAS = 1, PTC_TRI_PARTITIONED = 2
  integer(c_int), parameter, public :: PTC_POISSON_CHUNK_ROWS = 1, PTC_POISSON_SLAB_PLANES = 2
  integer(c_int), parameter, public :: PTC_POISSON_TRI_LINES = 3, PTC_POISSON_FAST_ZTILE = 4
  integer(c_int), parameter, public :: PTC_POISSON_TRI_LANES = 5, PTC_POISSON_TRI_COMPLEX = 6
  integer(c_int), parameter, public :: PTC_POISSON_STREAMING = 7

  interface
    integer(c_int) function ptc_create(grid, prm, yp, ctx) bind(C, name="ptc_create")
      import :: c_int, c_double, c_ptr, ptc_grid_t, ptc_params_t
      type(ptc_grid_t), intent(in) :: grid
      type(ptc_params_t), intent(in) :: prm
      real(c_double), intent(in) :: yp(*)
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function ptc_destroy(ctx) bind(C, name="ptc_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function ptc_last_error() bind(C, name="ptc_last_error")
      import :: c_ptr
    end function
    integer(c_int) function ptc_upload(ctx, field, host, layout) bind(C, name="ptc_upload")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: field, layout
      real(c_double), intent(in) :: host(*)
    end function
    integer(c_int) function ptc_download(ctx, field, host, layout) bind(C, name="ptc_download")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: field, layout
      real(c_double), intent(out) :: host(*)
    end function
    integer(c_int) function ptc_synchronize(ctx) bind(C, name="ptc_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ptc_init_poisson_solver(ctx, src, dst) bind(C, name="ptc_init_poisson_solver")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: src, dst
    end function
    integer(c_int) function ptc_clc_ppe(ctx, src, dst) bind(C, name="ptc_clc_ppe")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: src, dst
    end function
    integer(c_int) function ptc_destroy_poisson_plan(ctx) bind(C, name="ptc_destroy_poisson_plan")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ptc_set_bc_and_update_halo(ctx, ux, uy, uz, uy_ym) &
        bind(C, name="ptc_set_bc_and_update_halo")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ux, uy, uz, uy_ym
    end function
    integer(c_int) function ptc_set_bc_and_update_halo_pr(ctx, p) bind(C, name="ptc_set_bc_and_update_halo_pr")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: p
    end function
    integer(c_int) function ptc_set_bc_and_update_halo_scalar(ctx, s) &
        bind(C, name="ptc_set_bc_and_update_halo_scalar")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: s
    end function
    integer(c_int) function ptc_channel_iterate(ctx, nsteps, mon) bind(C, name="ptc_channel_iterate")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
      real(c_double), intent(out) :: mon(7)
    end function
    integer(c_int) function ptc_calc_ux_aver(ctx, ux, mean) bind(C, name="ptc_calc_ux_aver")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ux
      real(c_double), intent(out) :: mean
    end function
    integer(c_int) function ptc_check_divergence(ctx, ux, uy, uz, divmax) bind(C, name="ptc_check_divergence")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ux, uy, uz
      real(c_double), intent(out) :: divmax
    end function
    ! m_Parameters / m_TScheme (src/m_Parameters.f90:215-230, src/m_TScheme.f90:25-532)
    integer(c_int) function ptc_pm_coe_update(ctx, ns, dt) bind(C, name="ptc_pm_coe_update")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ns
      real(c_double), value :: dt
    end function
    integer(c_int) function ptc_set_coe(ctx, in9) bind(C, name="ptc_set_coe")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: in9(9)
    end function
    integer(c_int) function ptc_get_coe(ctx, out9) bind(C, name="ptc_get_coe")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out9(9)
    end function
    integer(c_int) function ptc_init_time_scheme(ctx) bind(C, name="ptc_init_time_scheme")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ptc_clc_rhs_xyz(ctx, ux, uy, uz, sc, rx, ry, rz, rc, hx, hy, hz, hc, p) &
        bind(C, name="ptc_clc_rhs_xyz")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ux, uy, uz, sc, rx, ry, rz, rc, hx, hy, hz, hc, p
    end function
    integer(c_int) function ptc_clc_u1hat(ctx, ux, rhsx) bind(C, name="ptc_clc_u1hat")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ux, rhsx
    end function
    integer(c_int) function ptc_clc_u2hat(ctx, uy, rhsy, duy_ym) bind(C, name="ptc_clc_u2hat")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: uy, rhsy, duy_ym
    end function
    integer(c_int) function ptc_clc_u3hat(ctx, uz, rhsz) bind(C, name="ptc_clc_u3hat")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: uz, rhsz
    end function
    integer(c_int) function ptc_clc_scalar(ctx, sc, rhsc) bind(C, name="ptc_clc_scalar")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: sc, rhsc
    end function
    integer(c_int) function ptc_clc_prsrc(ctx, ux, uy, uz, prsrc, divmax) bind(C, name="ptc_clc_prsrc")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ux, uy, uz, prsrc
      real(c_double), intent(out) :: divmax
    end function
    integer(c_int) function ptc_fluid_vel_update(ctx, phi, ux, uy, uz) bind(C, name="ptc_fluid_vel_update")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: phi, ux, uy, uz
    end function
    integer(c_int) function ptc_pressure_update(ctx, p, phi) bind(C, name="ptc_pressure_update")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: p, phi
    end function
    ! m_BC_and_Halo / m_FlowCase (src/m_BC_and_Halo.f90:69-103, src/m_FlowCase.f90:112-122)
    integer(c_int) function ptc_set_bc_and_update_halo_for_prsrc(ctx, ux, uy, uz, uy_ym) &
        bind(C, name="ptc_set_bc_and_update_halo_for_prsrc")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ux, uy, uz, uy_ym
    end function
    integer(c_int) function ptc_update_uy_ym(ctx, uy_ym, duy_ym) bind(C, name="ptc_update_uy_ym")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: uy_ym, duy_ym
    end function
    ! m_Tools monitors (src/m_Tools.f90:24-110,256-301)
    integer(c_int) function ptc_calc_max_cfl(ctx, ux, uy, uz, uddxmax) bind(C, name="ptc_calc_max_cfl")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ux, uy, uz
      real(c_double), intent(out) :: uddxmax
    end function
    integer(c_int) function ptc_calc_vmax(ctx, ux, uy, uz, vmax) bind(C, name="ptc_calc_vmax")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ux, uy, uz
      real(c_double), intent(out) :: vmax(3)
    end function
    integer(c_int) function ptc_clc_mean_value(ctx, ux, sc, res) bind(C, name="ptc_clc_mean_value")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ux, sc
      real(c_double), intent(out) :: res(3)
    end function
    ! Poisson plan (the scheme choice of InitPoissonSolver, src/m_Poisson.f90:50-171)
    integer(c_int) function ptc_poisson_config(ctx, prec, tri, tile, sweeps) bind(C, name="ptc_poisson_config")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: prec, tri, tile, sweeps
    end function
    integer(c_int) function ptc_poisson_set(ctx, knob, val) bind(C, name="ptc_poisson_set")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: knob, val
    end function
    ! the last plan search, one line per timing (ptc_poisson_search_log)
    integer(c_int) function ptc_poisson_search_log(ctx, buf, cap, len) bind(C, name="ptc_poisson_search_log")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(out) :: buf
      integer(c_int), value :: cap
      integer(c_int), intent(out) :: len
    end function
    integer(c_int) function ptc_poisson_get_knob(ctx, knob, val) bind(C, name="ptc_poisson_get_knob")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: knob
      integer(c_int), intent(out) :: val
    end function
    integer(c_int) function ptc_clc_stat(ctx, ux, uy, uz, pressure, scalar) bind(C, name="ptc_clc_stat")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ux, uy, uz, pressure, scalar
    end function
    integer(c_int) function ptc_get_stat(ctx, sumstat, nstat, nfstime, prgradsum, reset) bind(C, name="ptc_get_stat")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: sumstat(*)
      integer(c_int), intent(out) :: nstat, nfstime
      real(c_double), intent(out) :: prgradsum
      integer(c_int), value :: reset
    end function
    ! InitStatVar's SpectraOptions / Spectra2DOpt and clcStat's spectra
    ! (src/m_FlowCase.f90:140-251, 430-805); the sums come back column-major as
    ! EnergySpecX/Z and EnergySpec2D, not yet divided by nSpectime
    integer(c_int) function ptc_init_spectra(ctx, clc_1d, jfor_lcs, clc_2d, ny_spec2d, iy_spec2d) &
        bind(C, name="ptc_init_spectra")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: clc_1d, jfor_lcs, clc_2d, ny_spec2d
      integer(c_int), intent(in) :: iy_spec2d(*)
    end function
    integer(c_int) function ptc_clc_spectra(ctx, ux, uy, uz, pressure, scalar) bind(C, name="ptc_clc_spectra")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ux, uy, uz, pressure, scalar
    end function
    integer(c_int) function ptc_get_spectra(ctx, specx, specz, spec2d, nspectime, reset) bind(C, name="ptc_get_spectra")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: specx, specz, spec2d
      integer(c_int), intent(out) :: nspectime
      integer(c_int), value :: reset
    end function
    integer(c_int) function ptc_visu_field(ctx, which, iskip, jskip, kskip, real_bytes, host_out, count) &
        bind(C, name="ptc_visu_field")
      import :: c_int, c_ptr, c_long_long
      type(c_ptr), value :: ctx
      integer(c_int), value :: which, iskip, jskip, kskip, real_bytes
      type(c_ptr), value :: host_out
      integer(c_long_long), value :: count
    end function
    integer(c_int) function ptc_poisson_get_config(ctx, out10) bind(C, name="ptc_poisson_get_config")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: out10(10)
    end function
    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  function ptc_last_error_f() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = ptc_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    enddo
  end function ptc_last_error_f

  ! abort on a nonzero code, as CheckForError(ErrT_Abort, ...) does
  subroutine gpu_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    if (rc /= 0) then
      write(*, '(A,A,A,I0,A,A)') 'ParaTC-GPU error in ', where, ': code ', rc, ': ', ptc_last_error_f()
      error stop 1
    endif
  end subroutine gpu_check

  subroutine gpu_put(ctx, field, arr, layout)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: field, layout
    real(c_double), intent(in) :: arr(*)
    call gpu_check(ptc_upload(ctx, field, arr, layout), 'ptc_upload')
  end subroutine gpu_put

  subroutine gpu_get(ctx, field, arr, layout)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: field, layout
    real(c_double), intent(out) :: arr(*)
    call gpu_check(ptc_download(ctx, field, arr, layout), 'ptc_download')
  end subroutine gpu_get

  ! clcPPE(prsrc, prphiHalo) on host arrays (src/m_Poisson.f90:352-452): prsrc is
  ! interior, prphiHalo haloed; only the interior of prphiHalo is meaningful.
  subroutine gpu_clcPPE(ctx, prsrc, prphiHalo)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(in) :: prsrc(*)
    real(c_double), intent(inout) :: prphiHalo(*)
    call gpu_put(ctx, PTC_F_ARR1, prsrc, PTC_L_INTERIOR)
    call gpu_check(ptc_clc_ppe(ctx, PTC_F_ARR1, PTC_F_HALO), 'clcPPE')
    call gpu_get(ctx, PTC_F_HALO, prphiHalo, PTC_L_HALO)
  end subroutine gpu_clcPPE

end module m_paratc_gpu
