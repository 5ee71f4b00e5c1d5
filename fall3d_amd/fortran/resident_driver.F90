!***************************************************************
!> Resident driver: the ADS part of FALL3D's time loop with the data kept
!> on the GPU, through the fall3d_ads_c interfaces: grid, meteo interval,
!> source and initial state uploaded once, then nsteps calls of
!> fall3d_ads_step (source, x/y sweeps in iiter-parity order, z sweep with
!> deposition; Sources/mod_F3D.F90:1437-1491), a snapshot taken halfway and
!> restored into a second context that finishes the run, and the state,
!> fluxes and budget read back. Input/output are raw native-endian streams
!> written/read by tests/test_fortran_resident.py:
!>   in : int32 nx,ny,nz,nb,limiter,tm,nsteps,state ; real64 dt,stime ;
!>        dX1_p dX1_b dX2_p dX2_b dX3_p dX3_b u v w1 w2 k1 k2 k3 vs s c acum
!>   out: c acum flux(W,E,S,N,D,U) budget(12) ; then c acum of the restored run
!***************************************************************
program resident_driver
  use iso_c_binding
  use fall3d_ads_c
  implicit none
  integer(4) :: hdr(8)
  integer(c_int) :: nx, ny, nz, nb, limiter, tm, nsteps, state, iiter, rc, half
  real(c_double) :: dt, stime, flux(6), budget(FALL3D_ADS_BUDGET_LEN), t_rst
  integer(c_long_long) :: it_rst
  integer(c_size_t) :: nbytes
  character(len=1024) :: fin, fout
  real(c_double), allocatable, target :: dX1_p(:), dX1_b(:), dX2_p(:), dX2_b(:), dX3_p(:), dX3_b(:)
  real(c_double), allocatable, target :: u(:,:,:), v(:,:,:), w1(:,:,:), w2(:,:,:)
  real(c_double), allocatable, target :: k1(:,:,:), k2(:,:,:), k3(:,:,:), vs(:,:,:,:), s(:,:,:,:)
  real(c_double), allocatable, target :: c(:,:,:,:), acum(:,:,:), c2(:,:,:,:), acum2(:,:,:)
  integer(c_int8_t), allocatable, target :: blob(:)
  type(c_ptr) :: ctx, ctx2
  integer :: un

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=un, file=trim(fin), access='stream', form='unformatted', status='old')
  read(un) hdr
  nx = hdr(1); ny = hdr(2); nz = hdr(3); nb = hdr(4)
  limiter = hdr(5); tm = hdr(6); nsteps = hdr(7); state = hdr(8)
  read(un) dt, stime
  allocate(dX1_p(-1:nx+2), dX1_b(0:nx+2), dX2_p(-1:ny+2), dX2_b(0:ny+2), dX3_p(-1:nz+2), dX3_b(0:nz+2))
  allocate(u(0:nx+2, ny, nz), v(0:ny+2, nx, nz), w1(0:nz+2, nx, ny), w2(0:nz+2, nx, ny))
  allocate(k1(-1:nx+2, ny, nz), k2(-1:ny+2, nx, nz), k3(-1:nz+2, nx, ny), vs(nx, ny, 0:nz+2, nb))
  allocate(s(nx, ny, nz, nb), c(-1:nx+2, -1:ny+2, -1:nz+2, nb), acum(-1:nx+2, -1:ny+2, nb))
  allocate(c2, mold=c)
  allocate(acum2, mold=acum)
  read(un) dX1_p, dX1_b, dX2_p, dX2_b, dX3_p, dX3_b
  read(un) u, v, w1, w2, k1, k2, k3, vs, s, c, acum
  close(un)

  rc = fall3d_ads_initialize(ctx, nx, ny, nz, nb, limiter, tm, 8_c_int, state)
  call check(rc, ctx, 'initialize')
  call check(fall3d_ads_set_grid(ctx, c_loc(dX1_p), c_loc(dX1_b), c_loc(dX2_p), c_loc(dX2_b), c_loc(dX3_p), &
                                 c_loc(dX3_b)), ctx, 'set_grid')
  ! steady horizontal winds (u1 = u2, v1 = v2); w interpolated at stime
  call check(fall3d_ads_set_meteo(ctx, c_loc(u), c_loc(u), c_loc(v), c_loc(v), c_loc(w1), c_loc(w2), c_loc(k1), &
                                  c_loc(k2), c_loc(k3), c_loc(vs)), ctx, 'set_meteo')
  call check(fall3d_ads_set_source(ctx, c_loc(s)), ctx, 'set_source')
  call check(fall3d_ads_set_state(ctx, c_loc(c), c_loc(acum)), ctx, 'set_state')

  half = nsteps / 2
  do iiter = 0, half - 1
     call check(fall3d_ads_step(ctx, dt, stime, iiter), ctx, 'step')
  end do
  ! restart snapshot halfway: size query, then the snapshot itself
  nbytes = 0
  call check(fall3d_ads_snapshot(ctx, c_null_ptr, nbytes, real(half, c_double) * dt, int(half, c_long_long)), &
             ctx, 'snapshot size')
  allocate(blob(nbytes))
  call check(fall3d_ads_snapshot(ctx, c_loc(blob), nbytes, real(half, c_double) * dt, int(half, c_long_long)), &
             ctx, 'snapshot')
  do iiter = half, nsteps - 1
     call check(fall3d_ads_step(ctx, dt, stime, iiter), ctx, 'step')
  end do
  call check(fall3d_ads_get_state(ctx, c_loc(c), c_loc(acum)), ctx, 'get_state')
  call check(fall3d_ads_get_fluxes(ctx, flux), ctx, 'get_fluxes')
  call check(fall3d_ads_get_budget(ctx, budget), ctx, 'get_budget')

  ! a second context resumes from the snapshot and finishes the run
  rc = fall3d_ads_initialize(ctx2, nx, ny, nz, nb, limiter, tm, 8_c_int, state)
  call check(rc, ctx2, 'initialize 2')
  call check(fall3d_ads_set_grid(ctx2, c_loc(dX1_p), c_loc(dX1_b), c_loc(dX2_p), c_loc(dX2_b), c_loc(dX3_p), &
                                 c_loc(dX3_b)), ctx2, 'set_grid 2')
  call check(fall3d_ads_set_meteo(ctx2, c_loc(u), c_loc(u), c_loc(v), c_loc(v), c_loc(w1), c_loc(w2), c_loc(k1), &
                                  c_loc(k2), c_loc(k3), c_loc(vs)), ctx2, 'set_meteo 2')
  call check(fall3d_ads_set_source(ctx2, c_loc(s)), ctx2, 'set_source 2')
  call check(fall3d_ads_restore(ctx2, c_loc(blob), nbytes, t_rst, it_rst), ctx2, 'restore')
  do iiter = int(it_rst, c_int), nsteps - 1
     call check(fall3d_ads_step(ctx2, dt, stime, iiter), ctx2, 'step 2')
  end do
  call check(fall3d_ads_get_state(ctx2, c_loc(c2), c_loc(acum2)), ctx2, 'get_state 2')
  call check(fall3d_ads_release(ctx2), c_null_ptr, 'release 2')
  call check(fall3d_ads_release(ctx), c_null_ptr, 'release')

  open(newunit=un, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(un) c, acum, flux, budget
  write(un) c2, acum2
  close(un)
  write(*,'(a,i0,a,i0)') 'resident_driver: ', nsteps, ' steps done, restored at ', it_rst

contains

  subroutine check(rc, ctx, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: what
    if (rc /= 0) then
       if (c_associated(ctx)) then
          write(*,'(a,a,a,i0,a,a)') 'resident_driver: ', what, ' failed (', rc, '): ', fall3d_ads_error(ctx)
       else
          write(*,'(a,a,a,i0,a)') 'resident_driver: ', what, ' failed (', rc, ')'
       end if
       stop 1
    end if
  end subroutine check

end program resident_driver
