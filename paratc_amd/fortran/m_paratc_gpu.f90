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
  public :: gpu_check, gpu_put, gpu_get, gpu_clcPPE

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
