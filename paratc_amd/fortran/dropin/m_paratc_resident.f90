! m_paratc_resident.f90 -- the device context and host/device residency of the
! drop-in modules m_Poisson, m_TScheme, m_BC_and_Halo and m_Tools of this
! directory.
!
! Those modules keep the reference's procedure names and argument lists
! (src/m_Poisson.f90:22, src/m_TScheme.f90:14-19, src/m_BC_and_Halo.f90:13-17,
! src/m_Tools.f90:12-16), so a host built from the reference's other sources
! compiles its ChannelIterate (src/m_ChannelSystem.f90:80-145) against them
! unchanged.  Their arguments are the host's own arrays (m_Variables); this
! module maps each array, by its address, to a device-resident field of one
! HIP context and moves data only when the side holding the newer copy
! changes:
!   * the first time an array is seen it is uploaded (its host copy is the
!     newer one);
!   * an array a device routine writes becomes device-newer; later device
!     routines use the device copy and nothing crosses PCIe;
!   * paratc_to_host(a) downloads a device-newer array (call it before host
!     code reads the array: statistics, visualisation, restart files);
!   * paratc_host_wrote(a) marks an array the host has changed, so the next
!     device routine uploads it.
! An array is bound to the device field of the role it is first seen in
! (RealArr1 first seen as RhsX is the device's RealArr1, and stays so when the
! host passes it again as prsrc), or to a spare field when that role's field
! is bound to another array.  Array shapes select the layout: haloed
! (nx+4, ny+2, nz+4), interior (nx, ny, nz) or a wall plane (nx, nz).
!
! The context is created on the first call, from the host's m_Parameters
! (grid, physics, options) and m_MeshAndMetries (yp) -- the module variables
! the reference's routines read.  Failures abort as CheckForError(ErrT_Abort)
! does (src/m_LogInfo.f90:164-170).
module m_paratc_resident
  use iso_c_binding
  use m_paratc_gpu
  use m_Parameters
  use m_MeshAndMetries, only: yp
  implicit none
  private

  integer, parameter :: NF = 17                 ! device fields, PTC_F_UX .. PTC_F_DUY_YM
  integer, parameter :: HOST_NEWER = 0, IN_SYNC = 1, DEVICE_NEWER = 2

  type(c_ptr), save :: ctx = c_null_ptr
  type(c_ptr), save :: bound(0:NF - 1) = c_null_ptr    ! host array bound to each device field
  integer, save :: state(0:NF - 1) = HOST_NEWER
  integer(c_int), save :: layout(0:NF - 1) = PTC_L_HALO

  public :: paratc_ctx, paratc_in, paratc_out, paratc_inout, paratc_plane_in
  public :: paratc_to_host, paratc_host_wrote, paratc_push_coe, paratc_pull_prgrad
  public :: paratc_finalize, paratc_field_of

contains

  function paratc_ctx() result(c)
    type(c_ptr) :: c
    type(ptc_grid_t) :: g
    type(ptc_params_t) :: p
    if (.not. c_associated(ctx)) then
      g%nx = nxc; g%ny = nyc; g%nz = nzc; g%pad_ = 0
      g%xlx = xlx; g%yly = yly; g%zlz = zlz
      p%flow_type = FlowType
      p%is_ux_const = merge(1, 0, IsUxConst)
      p%ischeme = ischeme; p%icfl = iCFL; p%pad_ = 0
      p%ubulk = ubulk; p%xnu = xnu; p%ucrf = uCRF; p%dtmax = dtMax; p%cflc = CFLc
      p%gravity = gravity
      p%uxbc = uxBcValue; p%uybc = uyBcValue; p%uzbc = uzBcValue
      p%vel_limit = vel_limit; p%div_limit = div_limit
#ifdef ScalarFlow
      p%scalar = 1
      p%sdiff = SDiffCoe; p%gev = GravityEffVec; p%fvv = FallingVelVec
      p%sbc_opt = ScalarBcOption; p%sbc_val = ScalarBcValues
#else
      p%scalar = 0
      p%sdiff = xnu; p%gev = 0d0; p%fvv = 0d0
      p%sbc_opt = -2; p%sbc_val = 0d0
#endif
      call gpu_check(ptc_create(g, p, yp(1:nyc + 1), ctx), 'ptc_create')
    endif
    c = ctx
  end function paratc_ctx

  ! the layout of a host array of n elements
  function layout_of(n) result(l)
    integer(c_size_t), intent(in) :: n
    integer(c_int) :: l
    if (n == int(nxc + 4, c_size_t) * (nyc + 2) * (nzc + 4)) then
      l = PTC_L_HALO
    elseif (n == int(nxc, c_size_t) * nyc * nzc) then
      l = PTC_L_INTERIOR
    elseif (n == int(nxc, c_size_t) * nzc) then
      l = PTC_L_PLANE
    else
      write(*, '(A,I0)') 'ParaTC-GPU drop-in: an array of unexpected size ', n
      error stop 1
    endif
  end function layout_of

  ! the device field bound to the host array at addr (binding it on first
  ! sight: to `role`, else to a free spare field)
  function paratc_field_of(addr, n, role) result(f)
    type(c_ptr), intent(in) :: addr
    integer(c_size_t), intent(in) :: n
    integer(c_int), intent(in) :: role
    integer(c_int) :: f
    integer :: q
    do q = 0, NF - 1
      if (c_associated(bound(q), addr)) then
        f = q
        return
      endif
    enddo
    f = -1
    if (.not. c_associated(bound(role))) then
      f = role
    elseif (.not. c_associated(bound(PTC_F_WORK1))) then
      f = PTC_F_WORK1
    elseif (.not. c_associated(bound(PTC_F_WORK2))) then
      f = PTC_F_WORK2
    else
      write(*, '(A,I0)') 'ParaTC-GPU drop-in: no device field left for an array in role ', role
      error stop 1
    endif
    bound(f) = addr
    layout(f) = layout_of(n)
    state(f) = HOST_NEWER
  end function paratc_field_of

  ! a device routine reads the array: upload it if the host copy is newer
  function paratc_in(a, role) result(f)
    real(c_double), dimension(:, :, :), intent(in), target, contiguous :: a
    integer(c_int), intent(in) :: role
    integer(c_int) :: f
    type(c_ptr) :: c
    c = paratc_ctx()
    f = paratc_field_of(c_loc(a), size(a, kind=c_size_t), role)
    if (state(f) == HOST_NEWER) then
      call gpu_put(c, f, a, layout(f))
      state(f) = IN_SYNC
    endif
  end function paratc_in

  function paratc_plane_in(a, role) result(f)
    real(c_double), dimension(:, :), intent(in), target, contiguous :: a
    integer(c_int), intent(in) :: role
    integer(c_int) :: f
    type(c_ptr) :: c
    c = paratc_ctx()
    f = paratc_field_of(c_loc(a), size(a, kind=c_size_t), role)
    if (state(f) == HOST_NEWER) then
      call gpu_put(c, f, a, layout(f))
      state(f) = IN_SYNC
    endif
  end function paratc_plane_in

  ! a device routine writes the whole array without reading it
  function paratc_out(a, role) result(f)
    real(c_double), dimension(:, :, :), intent(in), target, contiguous :: a
    integer(c_int), intent(in) :: role
    integer(c_int) :: f
    type(c_ptr) :: c
    c = paratc_ctx()
    f = paratc_field_of(c_loc(a), size(a, kind=c_size_t), role)
    state(f) = DEVICE_NEWER
  end function paratc_out

  ! a device routine reads and writes the array
  function paratc_inout(a, role) result(f)
    real(c_double), dimension(:, :, :), intent(in), target, contiguous :: a
    integer(c_int), intent(in) :: role
    integer(c_int) :: f
    f = paratc_in(a, role)
    state(f) = DEVICE_NEWER
  end function paratc_inout

  ! the host is about to read the array: download it if the device copy is newer
  subroutine paratc_to_host(a)
    real(c_double), dimension(:, :, :), intent(inout), target, contiguous :: a
    integer :: q
    do q = 0, NF - 1
      if (c_associated(bound(q), c_loc(a))) then
        if (state(q) == DEVICE_NEWER) then
          call gpu_get(ctx, int(q, c_int), a, layout(q))
          state(q) = IN_SYNC
        endif
        return
      endif
    enddo
  end subroutine paratc_to_host

  ! the host has changed the array: the next device routine uploads it
  subroutine paratc_host_wrote(a)
    real(c_double), dimension(:, :, :), intent(in), target, contiguous :: a
    integer :: q
    do q = 0, NF - 1
      if (c_associated(bound(q), c_loc(a))) state(q) = HOST_NEWER
    enddo
  end subroutine paratc_host_wrote

  ! the substep coefficients of PMcoeUpdate (src/m_Parameters.f90:215-230) and
  ! PrGradAver, as the host's m_Parameters holds them, to the device
  subroutine paratc_push_coe()
    real(c_double) :: coe(9)
    coe(1) = pmGamma; coe(2) = pmTheta; coe(3) = pmAlpha; coe(4) = pmAlphaC; coe(5) = pmBeta
#ifdef ScalarFlow
    coe(6) = pmBetaSc
#else
    coe(6) = 0d0
#endif
    coe(7) = dt; coe(8) = SimTime; coe(9) = PrGradAver
    call gpu_check(ptc_set_coe(paratc_ctx(), coe), 'ptc_set_coe')
  end subroutine paratc_push_coe

  ! PrGradAver as clcRhsXYZ accumulated it (src/m_TScheme.f90:284,298), back
  ! into the host's m_Parameters
  subroutine paratc_pull_prgrad()
    real(c_double) :: coe(9)
    call gpu_check(ptc_get_coe(paratc_ctx(), coe), 'ptc_get_coe')
    PrGradAver = coe(9)
  end subroutine paratc_pull_prgrad

  subroutine paratc_finalize()
    integer(c_int) :: rc
    if (c_associated(ctx)) rc = ptc_destroy(ctx)
    ctx = c_null_ptr
    bound = c_null_ptr
    state = HOST_NEWER
  end subroutine paratc_finalize

end module m_paratc_resident
