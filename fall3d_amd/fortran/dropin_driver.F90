!***************************************************************
!> Drop-in driver: runs nsteps of the ADS part of F3D_time_step
!> (Sources/mod_F3D.F90:1435-1491 -- velocity pointers from ADS_get_vcomps,
!> host source injection, x/y sweeps in iiter-parity order, z sweep with
!> ground accumulation) calling the replacement MODULE ADS exactly as
!> mod_F3D.F90 calls the reference one. Input/output are raw native-endian
!> streams written/read by tests/test_fortran_dropin.py:
!>   in : int32 nx,ny,nz,nb,limiter,tm,nsteps,state ; real64 dt,stime ;
!>        dX1_p dX1_b dX2_p dX2_b dX3_p dX3_b u v w1 w2 k1 k2 k3 vs s c acum
!>   out: c acum flux(W,E,S,N,D,U)
!***************************************************************
program dropin_driver
  use KindType
  use Domain
  use ADS
  implicit none
  integer(4) :: hdr(8)
  integer(ip) :: nx, ny, nz, nb, limiter, tm, nsteps, state, iiter, i, j, k, b
  real(8) :: dt8, stime8
  real(rp) :: dt, stime, vol
  real(rp) :: W_flux, E_flux, S_flux, N_flux, D_flux, U_flux
  character(len=1024) :: fin, fout
  type(ARAKAWA_C_GRID) :: GRID
  real(rp), allocatable :: u(:,:,:), v(:,:,:), w1(:,:,:), w2(:,:,:)
  real(rp), allocatable :: k1(:,:,:), k2(:,:,:), k3(:,:,:), vs(:,:,:,:), s(:,:,:,:)
  real(rp), allocatable :: c(:,:,:,:), acum(:,:,:)
  real(rp), pointer :: my_u(:,:,:), my_v(:,:,:), my_w(:,:,:)
  integer :: un

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=un, file=trim(fin), access='stream', form='unformatted', status='old')
  read(un) hdr
  nx = hdr(1); ny = hdr(2); nz = hdr(3); nb = hdr(4)
  limiter = hdr(5); tm = hdr(6); nsteps = hdr(7); state = hdr(8)
  read(un) dt8, stime8
  dt = real(dt8, rp); stime = real(stime8, rp)
  call domain_serial(nx, ny, nz)
  allocate(GRID%dX1_p(my_ips_2h:my_ipe_2h), GRID%dX1_b(my_ibs_1h:my_ibe_1h))
  allocate(GRID%dX2_p(my_jps_2h:my_jpe_2h), GRID%dX2_b(my_jbs_1h:my_jbe_1h))
  allocate(GRID%dX3_p(my_kps_2h:my_kpe_2h), GRID%dX3_b(my_kbs_1h:my_kbe_1h))
  allocate(u (my_ibs_1h:my_ibe_1h, my_jps:my_jpe, my_kps:my_kpe))
  allocate(v (my_jbs_1h:my_jbe_1h, my_ips:my_ipe, my_kps:my_kpe))
  allocate(w1(my_kbs_1h:my_kbe_1h, my_ips:my_ipe, my_jps:my_jpe))
  allocate(w2(my_kbs_1h:my_kbe_1h, my_ips:my_ipe, my_jps:my_jpe))
  allocate(k1(my_ips_2h:my_ipe_2h, my_jps:my_jpe, my_kps:my_kpe))
  allocate(k2(my_jps_2h:my_jpe_2h, my_ips:my_ipe, my_kps:my_kpe))
  allocate(k3(my_kps_2h:my_kpe_2h, my_ips:my_ipe, my_jps:my_jpe))
  allocate(vs(my_ips:my_ipe, my_jps:my_jpe, my_kbs_1h:my_kbe_1h, nb))
  allocate(s (my_ips:my_ipe, my_jps:my_jpe, my_kps:my_kpe, nb))
  allocate(c (my_ips_2h:my_ipe_2h, my_jps_2h:my_jpe_2h, my_kps_2h:my_kpe_2h, nb))
  allocate(acum(my_ips_2h:my_ipe_2h, my_jps_2h:my_jpe_2h, nb))
  read(un) GRID%dX1_p, GRID%dX1_b, GRID%dX2_p, GRID%dX2_b, GRID%dX3_p, GRID%dX3_b
  read(un) u, v, w1, w2, k1, k2, k3, vs, s, c, acum
  close(un)

  W_flux = 0.0_rp; E_flux = 0.0_rp; S_flux = 0.0_rp; N_flux = 0.0_rp; D_flux = 0.0_rp; U_flux = 0.0_rp
  call ADS_set_state_mode(state)
  call ADS_initialize(limiter, tm)
  do iiter = 0, nsteps - 1
     call ADS_get_vcomps(my_u, my_v, my_w)
     my_u(:,:,:) = u(:,:,:)
     my_v(:,:,:) = v(:,:,:)
     do k = my_kps, my_kpe
        do j = my_jps, my_jpe
           do i = my_ips, my_ipe
              vol = GRID%dX1_p(i) * GRID%dX2_p(j) * GRID%dX3_p(k)
              do b = 1, nb
                 c(i,j,k,b) = c(i,j,k,b) + dt * s(i,j,k,b) / vol
              end do
           end do
        end do
     end do
     if (mod(iiter, 2) == 0) then
        call ADS_solve_along_x(W_flux, E_flux, dt, c, k1, GRID, nb)
        call ADS_solve_along_y(S_flux, N_flux, dt, c, k2, GRID, nb)
     else
        call ADS_solve_along_y(S_flux, N_flux, dt, c, k2, GRID, nb)
        call ADS_solve_along_x(W_flux, E_flux, dt, c, k1, GRID, nb)
     end if
     call ADS_solve_along_z(D_flux, U_flux, dt, c, acum, k3, w1, w2, stime, vs, GRID, nb)
  end do
  call ADS_release()

  open(newunit=un, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(un) c, acum
  write(un) real(W_flux, 8), real(E_flux, 8), real(S_flux, 8), real(N_flux, 8), real(D_flux, 8), real(U_flux, 8)
  close(un)
  write(*,'(a,i0,a)') 'dropin_driver: ', nsteps, ' steps done'
end program dropin_driver
