! Drives the C ABI from Fortran through the diablo_hip module: reads a deck
! and a state written by the test (stream files), runs `nsteps` time steps,
! writes the state back.  Exercises the bind(C) layouts and the array
! wrappers the RK_CHAN_1 drop-in uses.
program drive_substep
  use diablo_hip
  implicit none
  type(dh_params) :: p
  character(len=1) :: raw(c_sizeof(p))
  type(c_ptr) :: ctx
  integer(c_int) :: rc
  integer :: ny, nz, nx, nxp, nth, nsteps, ts, u
  real(c_double) :: time, dt, t2, hb(3)
  integer(c_int32_t) :: ts2
  real(c_double), allocatable :: gy(:), gyf(:), dy(:), dyf(:), th(:,:,:,:)
  complex(c_double_complex), allocatable :: cu1(:,:,:), cu2(:,:,:), cu3(:,:,:), cp(:,:,:)
  complex(c_double_complex), allocatable :: cf1(:,:,:), cf2(:,:,:), cf3(:,:,:), cth(:,:,:,:), cfth(:,:,:,:)

  open(newunit=u, file='fortran_in.bin', access='stream', form='unformatted', status='old')
  ! raw C bytes including padding: unformatted I/O of the type itself would skip it
  read(u) raw
  p = transfer(raw, p)
  read(u) nsteps, ts, time, dt
  nx = p%nx; ny = p%ny; nz = p%nz; nth = max(p%n_th, 1); nxp = nx / 2
  allocate(gy(0:ny+1), gyf(0:ny+1), dy(0:ny+1), dyf(0:ny+1))
  allocate(cu1(0:nxp,0:nz+1,0:ny+1), cu2(0:nxp,0:nz+1,0:ny+1), cu3(0:nxp,0:nz+1,0:ny+1))
  allocate(cp(0:nxp,0:nz+1,0:ny+1), cf1(0:nxp,0:nz+1,0:ny+1), cf2(0:nxp,0:nz+1,0:ny+1))
  allocate(cf3(0:nxp,0:nz+1,0:ny+1), cth(0:nxp,0:nz+1,0:ny+1,nth), cfth(0:nxp,0:nz+1,0:ny+1,nth))
  allocate(th(0:nx+1,0:nz+1,0:ny+1,nth))
  read(u) gy, gyf, dy, dyf
  read(u) cu1, cu2, cu3, cp, cf1, cf2, cf3, cth, cfth, th
  close(u)
  write(*,'(a,i0)') 'sizeof(dh_params) = ', c_sizeof(p)

  rc = dh_create(p, ctx)
  if (rc /= 0) stop 'dh_create failed'
  call check(dh_set_grid(ctx, gy, gyf, dy, dyf), 'dh_set_grid')
  call check(dh_set_clock(ctx, time, dt, int(ts, c_int32_t)), 'dh_set_clock')
  call check(dh_upload_arrays(ctx, cu1, cu2, cu3, cp, cf1, cf2, cf3, cth, cfth, th), 'dh_upload')
  call check(dh_time_step(ctx, int(nsteps, c_int32_t)), 'dh_time_step')
  call check(dh_synchronize(ctx), 'dh_synchronize')
  call check(dh_download_arrays(ctx, cu1, cu2, cu3, cp, cf1, cf2, cf3, cth, cfth, th), 'dh_download')
  call check(dh_get_clock(ctx, t2, dt, hb, ts2), 'dh_get_clock')
  rc = dh_destroy(ctx)

  open(newunit=u, file='fortran_out.bin', access='stream', form='unformatted', status='replace')
  write(u) t2, dt, int(ts2)
  write(u) cu1, cu2, cu3, cp, cth, th
  close(u)
  write(*,'(a)') 'drive_substep: done'
contains
  subroutine check(r, what)
    integer(c_int), intent(in) :: r
    character(*), intent(in) :: what
    if (r /= 0) then
      write(*,*) what, ' failed: ', trim(dh_error_text(ctx))
      stop 1
    end if
  end subroutine check
end program drive_substep
