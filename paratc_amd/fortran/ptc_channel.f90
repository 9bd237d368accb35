! ptc_channel.f90 -- a Fortran host that advances a case on the GPU through
! m_paratc_gpu, reading and writing the reference's restart format.
!
! Usage: ptc_channel <control file>
! The control file is a namelist &ptc_run (grid, physics, step count, file
! names).  yp(1..ny+1) is read from a raw FP64 file (the host's clcYcoord
! result).  The restart files are raw native FP64 with no header, each field
! the full nx*ny*nz interior in Fortran order: ux(+uCRF), uy, uz, p
! (write_restart / read_restart, src/m_IOAndVisu.f90:408-512).
! With spec_out set it also writes one sample of clcStat's 1-D energy spectra
! of the final fields, as the reference's SpecX / SpecZ files
! (src/m_FlowCase.f90:430-805, 841-885), to spec_out//'SpecX' and 'SpecZ'.
program ptc_channel
  use iso_c_binding
  use m_paratc_gpu
  implicit none
  integer :: nx, ny, nz, nsteps, flow_type, is_ux_const, ischeme
  real(8) :: xlx, yly, zlz, xnu, ubulk, dtmax
  character(len=256) :: yp_file, restart_in, restart_out, ctl, spec_out
  integer :: jforlcs
  namelist /ptc_run/ nx, ny, nz, xlx, yly, zlz, xnu, ubulk, dtmax, nsteps, flow_type, &
                     is_ux_const, ischeme, yp_file, restart_in, restart_out, spec_out, jforlcs
  type(ptc_grid_t) :: grid
  type(ptc_params_t) :: prm
  type(c_ptr) :: ctx
  real(c_double), allocatable :: yp(:), ux(:,:,:), uy(:,:,:), uz(:,:,:), pr(:,:,:)
  real(c_double) :: mon(7), umean
  real(c_double), allocatable, target :: specx(:,:,:), specz(:,:,:)
  integer(c_int) :: nspec, norows(1)
  integer :: u, ios
  integer(8) :: nbytes

  call get_command_argument(1, ctl)
  flow_type = 1; is_ux_const = 1; ischeme = 3
  spec_out = ''; jforlcs = 1
  open(newunit=u, file=trim(ctl), status='old', action='read', iostat=ios)
  if (ios /= 0) then
    write(*,*) 'Cannot open file: ', trim(ctl)
    error stop 1
  endif
  read(u, nml=ptc_run)
  close(u)

  allocate(yp(ny + 1), ux(nx, ny, nz), uy(nx, ny, nz), uz(nx, ny, nz), pr(nx, ny, nz))
  open(newunit=u, file=trim(yp_file), access='stream', form='unformatted', status='old', action='read')
  read(u) yp
  close(u)
  ! read_restart: byte count check as src/m_IOAndVisu.f90:477-493
  nbytes = 8_8 * nx * ny * nz * 4
  open(newunit=u, file=trim(restart_in), access='stream', form='unformatted', status='old', action='read', iostat=ios)
  if (ios /= 0) then
    write(*,*) 'Cannot open file: ', trim(restart_in)
    error stop 1
  endif
  inquire(unit=u, size=nbytes)
  if (nbytes /= 8_8 * nx * ny * nz * 4) then
    write(*,*) 'file byte wrong2'
    error stop 1
  endif
  read(u) ux, uy, uz, pr
  close(u)

  grid%nx = nx; grid%ny = ny; grid%nz = nz; grid%pad_ = 0
  grid%xlx = xlx; grid%yly = yly; grid%zlz = zlz
  prm%flow_type = flow_type; prm%scalar = 0; prm%is_ux_const = is_ux_const
  prm%ischeme = ischeme; prm%icfl = 2; prm%sbc_opt = -2; prm%pad_ = 0
  prm%ubulk = ubulk; prm%xnu = xnu; prm%ucrf = 0d0; prm%dtmax = dtmax; prm%cflc = 1.5d0
  prm%gravity = 0d0; prm%uxbc = 0d0; prm%uybc = 0d0; prm%uzbc = 0d0
  prm%sdiff = xnu; prm%gev = 0d0; prm%fvv = 0d0; prm%sbc_val = 0d0
  prm%vel_limit = 0d0; prm%div_limit = 0d0

  call gpu_check(ptc_create(grid, prm, yp, ctx), 'ptc_create')
  call gpu_put(ctx, PTC_F_UX, ux, PTC_L_INTERIOR)
  call gpu_put(ctx, PTC_F_UY, uy, PTC_L_INTERIOR)
  call gpu_put(ctx, PTC_F_UZ, uz, PTC_L_INTERIOR)
  call gpu_put(ctx, PTC_F_PR, pr, PTC_L_INTERIOR)
  ! ChannelInitialize (src/m_ChannelSystem.f90:41,65-66)
  call gpu_check(ptc_init_poisson_solver(ctx, PTC_F_ARR1, PTC_F_HALO), 'InitPoissonSolver')
  call gpu_check(ptc_set_bc_and_update_halo(ctx, PTC_F_UX, PTC_F_UY, PTC_F_UZ, PTC_F_UY_YM), 'SetBC_and_UpdateHalo')
  call gpu_check(ptc_set_bc_and_update_halo_pr(ctx, PTC_F_PR), 'SetBC_and_UpdateHalo_pr')
  ! the time loop (src/main.f90:31-33)
  call gpu_check(ptc_channel_iterate(ctx, nsteps, mon), 'ChannelIterate')
  call gpu_check(ptc_calc_ux_aver(ctx, PTC_F_UX, umean), 'CalcUxAver')
  write(*, '(A,I0,A,ES12.5,A,ES12.5,A,3ES12.5)') 'steps ', nsteps, ' div ', mon(2), ' uxmean ', umean, &
        ' vmax ', mon(3:5)
  call gpu_get(ctx, PTC_F_UX, ux, PTC_L_INTERIOR)
  call gpu_get(ctx, PTC_F_UY, uy, PTC_L_INTERIOR)
  call gpu_get(ctx, PTC_F_UZ, uz, PTC_L_INTERIOR)
  call gpu_get(ctx, PTC_F_PR, pr, PTC_L_INTERIOR)
  if (len_trim(spec_out) > 0) then
    allocate(specx(nx/2 + 1, ny, 8), specz(nz/2 + 1, ny, 8))
    norows = 0
    call gpu_check(ptc_init_spectra(ctx, 1, jforlcs, 0, 0, norows), 'InitStatVar spectra')
    call gpu_check(ptc_clc_spectra(ctx, PTC_F_UX, PTC_F_UY, PTC_F_UZ, PTC_F_PR, -1), 'clcStat spectra')
    call gpu_check(ptc_get_spectra(ctx, c_loc(specx), c_loc(specz), c_null_ptr, nspec, 1), 'ptc_get_spectra')
    specx = specx / real(nspec, 8)
    specz = specz / real(nspec, 8)
    open(newunit=u, file=trim(spec_out)//'SpecX', access='stream', form='unformatted', status='replace', &
         action='write')
    write(u) specx
    close(u)
    open(newunit=u, file=trim(spec_out)//'SpecZ', access='stream', form='unformatted', status='replace', &
         action='write')
    write(u) specz
    close(u)
  endif
  call gpu_check(ptc_destroy(ctx), 'ptc_destroy')
  open(newunit=u, file=trim(restart_out), access='stream', form='unformatted', status='replace', action='write')
  write(u) ux, uy, uz, pr
  close(u)
end program ptc_channel
