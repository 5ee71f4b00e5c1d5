!***************************************************************
!>
!> MODULE ADS -- drop-in replacement of FALL3D's Sources/mod_ADS.F90 that
!> runs the advection-diffusion-sedimentation sweeps on an MI355X (gfx950)
!> through the C ABI of libfall3d_ads_hip.so (include/fall3d_ads.h).
!>
!> Same module name, same PUBLIC procedures, same dummy arguments, kinds
!> and intents as Sources/mod_ADS.F90:48-55, so Sources/mod_F3D.F90 compiles
!> and runs unchanged when this file replaces mod_ADS.F90 in the build.
!> It reads the same module globals as the reference (KindType: rp, ip,
!> ARAKAWA_C_GRID; Domain: my_ips..my_kbe_1h, my_*_proc).
!>
!> The GPU context needs the number of bins, which the reference learns
!> only at the first sweep, so it is created (or re-created when nbins
!> changes) at the first ADS_solve_along_* call. CB%u/v/w stay host arrays
!> owned by this module, as in the reference (:18-20), and are copied to
!> the context's staging arrays at each sweep. Failures of the C calls stop
!> the run with the library's message (the reference has no error dummy).
!>
!***************************************************************
MODULE ADS
  use KindType
  use Domain
  use iso_c_binding
  implicit none
  private
  !
  type ADS_CB_HIP
     real(rp), allocatable :: u(:,:,:)
     real(rp), allocatable :: v(:,:,:)
     real(rp), allocatable :: w(:,:,:)
     integer(ip) :: limiter       = 1
     integer(ip) :: time_marching = 2
     integer(ip) :: nbins         = 0
     type(c_ptr) :: ctx           = c_null_ptr
  end type ADS_CB_HIP
  !
  type(ADS_CB_HIP), TARGET, SAVE :: CB
  !
  ! device state: FP32 bulk (0) or compensated FP32 pair (1); the Fortran
  ! program may select it with ADS_set_state_mode before the first sweep
  integer(c_int), SAVE :: state_mode = 0_c_int
  !
  PUBLIC :: ADS_initialize
  PUBLIC :: ADS_get_vcomps
  PUBLIC :: ADS_update_horizontal_vcomps
  PUBLIC :: ADS_release
  PUBLIC :: ADS_solve_along_x
  PUBLIC :: ADS_solve_along_y
  PUBLIC :: ADS_solve_along_z
  PUBLIC :: ADS_solve_1D
  PUBLIC :: ADS_set_state_mode
  !
  interface
     integer(c_int) function f3d_initialize(ctx, nx, ny, nz, nbins, limiter, tm, kind, state) &
          bind(C, name="fall3d_ads_initialize")
       import :: c_ptr, c_int
       type(c_ptr)           :: ctx
       integer(c_int), value :: nx, ny, nz, nbins, limiter, tm, kind, state
     end function f3d_initialize
     integer(c_int) function f3d_release(ctx) bind(C, name="fall3d_ads_release")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
     end function f3d_release
     type(c_ptr) function f3d_last_error(ctx) bind(C, name="fall3d_ads_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function f3d_last_error
     integer(c_int) function f3d_set_grid(ctx, p1, b1, p2, b2, p3, b3) bind(C, name="fall3d_ads_set_grid")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, p1, b1, p2, b2, p3, b3
     end function f3d_set_grid
     integer(c_int) function f3d_get_vcomps(ctx, u, v, w) bind(C, name="fall3d_ads_get_vcomps")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       type(c_ptr)        :: u, v, w
     end function f3d_get_vcomps
     integer(c_int) function f3d_update_hv(ctx, u1, u2, v1, v2, stime) &
          bind(C, name="fall3d_ads_update_horizontal_vcomps")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: ctx, u1, u2, v1, v2
       real(c_double), value :: stime
     end function f3d_update_hv
     integer(c_int) function f3d_solve_x(ctx, fb, fe, dt, c, k) bind(C, name="fall3d_ads_solve_along_x")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: ctx, c, k
       real(c_double)        :: fb, fe
       real(c_double), value :: dt
     end function f3d_solve_x
     integer(c_int) function f3d_solve_y(ctx, fb, fe, dt, c, k) bind(C, name="fall3d_ads_solve_along_y")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: ctx, c, k
       real(c_double)        :: fb, fe
       real(c_double), value :: dt
     end function f3d_solve_y
     integer(c_int) function f3d_solve_z(ctx, fb, fe, dt, c, acum, k, w1, w2, stime, vs) &
          bind(C, name="fall3d_ads_solve_along_z")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: ctx, c, acum, k, w1, w2, vs
       real(c_double)        :: fb, fe
       real(c_double), value :: dt, stime
     end function f3d_solve_z
     integer(c_int) function f3d_solve_1d(ctx, c, dxp, dxb, u, k, dt, bvalue, flux, bcode, ips, ipe, limiter) &
          bind(C, name="fall3d_ads_solve_1d")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: ctx, c, dxp, dxb, u, k
       real(c_double), value :: dt
       real(c_double)        :: bvalue(2), flux(2)
       integer(c_int)        :: bcode(2)
       integer(c_int), value :: ips, ipe, limiter
     end function f3d_solve_1d
  end interface
  !
CONTAINS
  !
  subroutine check(rc, what)
    integer(c_int),   intent(IN) :: rc
    character(len=*), intent(IN) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: txt
    integer :: i
    if (rc == 0) return
    txt = ' '
    call c_f_pointer(f3d_last_error(CB%ctx), msg, [512])
    do i = 1, 512
       if (msg(i) == c_null_char) exit
       txt(i:i) = msg(i)
    end do
    write(*,'(a,a,a,i0,a,a)') 'ADS (HIP): ', what, ' failed (', rc, '): ', trim(txt)
    error stop 1
  end subroutine check
  !
  !> Select the device state (0 FP32, 1 compensated FP32 pair) of contexts
  !> created after this call. Not part of the reference interface.
  subroutine ADS_set_state_mode(mode)
    integer(ip), intent(IN) :: mode
    state_mode = int(mode, c_int)
  end subroutine ADS_set_state_mode
  !
  subroutine ensure_ctx(nbins)
    integer(ip), intent(IN) :: nbins
    integer(c_int) :: rc
    if (c_associated(CB%ctx) .and. CB%nbins == nbins) return
    if (c_associated(CB%ctx)) rc = f3d_release(CB%ctx)
    CB%ctx = c_null_ptr
    rc = f3d_initialize(CB%ctx, int(my_ipe-my_ips+1, c_int), int(my_jpe-my_jps+1, c_int), &
         int(my_kpe-my_kps+1, c_int), int(nbins, c_int), int(CB%limiter, c_int), &
         int(CB%time_marching, c_int), int(rp, c_int), state_mode)
    call check(rc, 'fall3d_ads_initialize')
    CB%nbins = nbins
  end subroutine ensure_ctx
  !
  subroutine push_grid(MY_GRID)
    type(ARAKAWA_C_GRID), target, intent(IN) :: MY_GRID
    integer(c_int) :: rc
    rc = f3d_set_grid(CB%ctx, c_loc(MY_GRID%dX1_p), c_loc(MY_GRID%dX1_b), c_loc(MY_GRID%dX2_p), &
         c_loc(MY_GRID%dX2_b), c_loc(MY_GRID%dX3_p), c_loc(MY_GRID%dX3_b))
    call check(rc, 'fall3d_ads_set_grid')
  end subroutine push_grid
  !
  !> copy CB%u / CB%v into the context's staging arrays (what the C sweep reads)
  subroutine push_vcomps(which)
    integer, intent(IN) :: which
    type(c_ptr) :: pu, pv, pw
    real(rp), pointer :: s(:)
    integer(c_int) :: rc
    rc = f3d_get_vcomps(CB%ctx, pu, pv, pw)
    call check(rc, 'fall3d_ads_get_vcomps')
    if (which == 1) then
       call c_f_pointer(pu, s, [size(CB%u)])
       s = reshape(CB%u, [size(CB%u)])
    else
       call c_f_pointer(pv, s, [size(CB%v)])
       s = reshape(CB%v, [size(CB%v)])
    end if
  end subroutine push_vcomps
  !
  !-----------------------------------
  !    subroutine ADS_initialize   (Sources/mod_ADS.F90:98)
  !-----------------------------------
  subroutine ADS_initialize(limitertype,time_marching)
    implicit none
    integer(ip), intent(IN) :: limitertype
    integer(ip), intent(IN) :: time_marching
    allocate(CB%u(my_ibs_1h:my_ibe_1h,my_jps   :my_jpe   ,my_kps   :my_kpe   ))
    allocate(CB%v(my_jbs_1h:my_jbe_1h,my_ips   :my_ipe   ,my_kps   :my_kpe   ))
    allocate(CB%w(my_kbs_1h:my_kbe_1h,my_ips   :my_ipe   ,my_jps   :my_jpe   ))
    CB%u = 0.0_rp
    CB%v = 0.0_rp
    CB%w = 0.0_rp
    CB%limiter       = limitertype
    CB%time_marching = time_marching
    CB%nbins         = 0
  end subroutine ADS_initialize
  !
  !----------------------------------------
  !    subroutine ADS_get_vcomps   (Sources/mod_ADS.F90:166)
  !----------------------------------------
  subroutine ADS_get_vcomps(my_u,my_v,my_w)
    implicit none
    real(rp),pointer,intent(INOUT) :: my_u(:,:,:)
    real(rp),pointer,intent(INOUT) :: my_v(:,:,:)
    real(rp),pointer,intent(INOUT) :: my_w(:,:,:)
    my_u => CB%u
    my_v => CB%v
    my_w => CB%w
  end subroutine ADS_get_vcomps
  !
  !----------------------------------------
  !    subroutine ADS_update_horizontal_vcomps   (Sources/mod_ADS.F90:187)
  !----------------------------------------
  subroutine ADS_update_horizontal_vcomps(my_u1,my_u2,my_v1,my_v2,stime)
    implicit none
    real(rp), target, intent(IN) :: my_u1 (my_ibs_1h:my_ibe_1h,my_jps   :my_jpe   ,my_kps   :my_kpe   )
    real(rp), target, intent(IN) :: my_u2 (my_ibs_1h:my_ibe_1h,my_jps   :my_jpe   ,my_kps   :my_kpe   )
    real(rp), target, intent(IN) :: my_v1 (my_jbs_1h:my_jbe_1h,my_ips   :my_ipe   ,my_kps   :my_kpe   )
    real(rp), target, intent(IN) :: my_v2 (my_jbs_1h:my_jbe_1h,my_ips   :my_ipe   ,my_kps   :my_kpe   )
    real(rp),         intent(IN) :: stime
    type(c_ptr) :: pu, pv, pw
    real(rp), pointer :: s(:)
    integer(c_int) :: rc
    call ensure_ctx(max(CB%nbins, 1_ip))
    rc = f3d_update_hv(CB%ctx, c_loc(my_u1), c_loc(my_u2), c_loc(my_v1), c_loc(my_v2), real(stime, c_double))
    call check(rc, 'fall3d_ads_update_horizontal_vcomps')
    rc = f3d_get_vcomps(CB%ctx, pu, pv, pw)
    call check(rc, 'fall3d_ads_get_vcomps')
    call c_f_pointer(pu, s, [size(CB%u)])
    CB%u = reshape(s, shape(CB%u))
    call c_f_pointer(pv, s, [size(CB%v)])
    CB%v = reshape(s, shape(CB%v))
  end subroutine ADS_update_horizontal_vcomps
  !
  !-----------------------------------
  !    subroutine ADS_release   (Sources/mod_ADS.F90:208)
  !-----------------------------------
  subroutine ADS_release()
    implicit none
    integer(c_int) :: rc
    if (c_associated(CB%ctx)) then
       rc = f3d_release(CB%ctx)
       CB%ctx = c_null_ptr
    end if
    CB%nbins = 0
    if (allocated(CB%u)) deallocate(CB%u)
    if (allocated(CB%v)) deallocate(CB%v)
    if (allocated(CB%w)) deallocate(CB%w)
  end subroutine ADS_release
  !
  !-----------------------------------
  !    subroutine ADS_solve_along_x   (Sources/mod_ADS.F90:233)
  !-----------------------------------
  subroutine ADS_solve_along_x(my_W_flux,my_E_flux,dt,my_c,my_k1,MY_GRID,nbins)
    implicit none
    integer(ip),         intent(IN   ) :: nbins
    real(rp),            intent(INOUT) :: my_W_flux
    real(rp),            intent(INOUT) :: my_E_flux
    real(rp),            intent(IN   ) :: dt
    real(rp),target,     intent(INOUT) :: my_c  (my_ips_2h:my_ipe_2h,my_jps_2h:my_jpe_2h,my_kps_2h:my_kpe_2h,nbins)
    real(rp),target,     intent(IN   ) :: my_k1 (my_ips_2h:my_ipe_2h,my_jps   :my_jpe   ,my_kps   :my_kpe   )
    type(ARAKAWA_C_GRID),target,intent(IN   ) :: MY_GRID
    real(c_double) :: fb, fe
    integer(c_int) :: rc
    call ensure_ctx(nbins)
    call push_grid(MY_GRID)
    call push_vcomps(1)
    fb = real(my_W_flux, c_double)
    fe = real(my_E_flux, c_double)
    rc = f3d_solve_x(CB%ctx, fb, fe, real(dt, c_double), c_loc(my_c), c_loc(my_k1))
    call check(rc, 'fall3d_ads_solve_along_x')
    my_W_flux = real(fb, rp)
    my_E_flux = real(fe, rp)
  end subroutine ADS_solve_along_x
  !
  !-----------------------------------
  !    subroutine ADS_solve_along_y   (Sources/mod_ADS.F90:422)
  !-----------------------------------
  subroutine ADS_solve_along_y(my_S_flux,my_N_flux,dt,my_c,my_k2,MY_GRID,nbins)
    implicit none
    integer(ip),         intent(IN   ) :: nbins
    real(rp),            intent(INOUT) :: my_S_flux
    real(rp),            intent(INOUT) :: my_N_flux
    real(rp),            intent(IN   ) :: dt
    real(rp),target,     intent(INOUT) :: my_c  (my_ips_2h:my_ipe_2h,my_jps_2h:my_jpe_2h,my_kps_2h:my_kpe_2h,nbins)
    real(rp),target,     intent(IN   ) :: my_k2 (my_jps_2h:my_jpe_2h,my_ips   :my_ipe   ,my_kps   :my_kpe   )
    type(ARAKAWA_C_GRID),target,intent(IN   ) :: MY_GRID
    real(c_double) :: fb, fe
    integer(c_int) :: rc
    call ensure_ctx(nbins)
    call push_grid(MY_GRID)
    call push_vcomps(2)
    fb = real(my_S_flux, c_double)
    fe = real(my_N_flux, c_double)
    rc = f3d_solve_y(CB%ctx, fb, fe, real(dt, c_double), c_loc(my_c), c_loc(my_k2))
    call check(rc, 'fall3d_ads_solve_along_y')
    my_S_flux = real(fb, rp)
    my_N_flux = real(fe, rp)
  end subroutine ADS_solve_along_y
  !
  !-----------------------------------
  !    subroutine ADS_solve_along_z   (Sources/mod_ADS.F90:612)
  !-----------------------------------
  subroutine ADS_solve_along_z(my_D_flux,my_U_flux,dt,my_c,my_acum,my_k3,my_w1,my_w2,stime,my_vs,MY_GRID,nbins)
    implicit none
    integer(ip),         intent(IN   ) :: nbins
    real(rp),            intent(INOUT) :: my_D_flux
    real(rp),            intent(INOUT) :: my_U_flux
    real(rp),            intent(IN   ) :: dt
    real(rp),target     ,intent(INOUT) :: my_c   (my_ips_2h:my_ipe_2h,my_jps_2h:my_jpe_2h,my_kps_2h:my_kpe_2h,nbins)
    real(rp),target     ,intent(INOUT) :: my_acum(my_ips_2h:my_ipe_2h,my_jps_2h:my_jpe_2h,nbins)
    real(rp),target     ,intent(IN   ) :: my_k3  (my_kps_2h:my_kpe_2h,my_ips   :my_ipe   ,my_jps   :my_jpe)
    real(rp),target     ,intent(IN   ) :: my_w1  (my_kbs_1h:my_kbe_1h,my_ips   :my_ipe   ,my_jps   :my_jpe)
    real(rp),target     ,intent(IN   ) :: my_w2  (my_kbs_1h:my_kbe_1h,my_ips   :my_ipe   ,my_jps   :my_jpe)
    real(rp),            intent(IN   ) :: stime
    real(rp),target     ,intent(IN   ) :: my_vs  (my_ips   :my_ipe   ,my_jps   :my_jpe   ,my_kbs_1h:my_kbe_1h,nbins)
    type(ARAKAWA_C_GRID),target,intent(IN   ) :: MY_GRID
    real(c_double) :: fb, fe
    integer(c_int) :: rc
    call ensure_ctx(nbins)
    call push_grid(MY_GRID)
    fb = real(my_D_flux, c_double)
    fe = real(my_U_flux, c_double)
    rc = f3d_solve_z(CB%ctx, fb, fe, real(dt, c_double), c_loc(my_c), c_loc(my_acum), c_loc(my_k3), &
         c_loc(my_w1), c_loc(my_w2), real(stime, c_double), c_loc(my_vs))
    call check(rc, 'fall3d_ads_solve_along_z')
    my_D_flux = real(fb, rp)
    my_U_flux = real(fe, rp)
  end subroutine ADS_solve_along_z
  !
  !-----------------------------------
  !    subroutine ADS_solve_1D   (Sources/mod_ADS.F90:820)
  !-----------------------------------
  subroutine ADS_solve_1D(c,dxp,dxb,u,k,dt,bvalue,flux,bcode,ips,ipe,ibs,ibe,limiter,idime)
    implicit none
    integer(ip), intent(in)    :: ips
    integer(ip), intent(in)    :: ipe
    integer(ip), intent(in)    :: ibs
    integer(ip), intent(in)    :: ibe
    integer(ip), intent(in)    :: bcode(2)
    integer(ip), intent(in)    :: limiter
    integer(ip), intent(in)    :: idime
    real   (rp), target, intent(inout) :: c      (ips-2:ipe+2)
    real   (rp), target, intent(in)    :: dxp    (ips-2:ipe+2)
    real   (rp), target, intent(in)    :: dxb    (ibs-1:ibe+1)
    real   (rp), target, intent(in)    :: u      (ibs-1:ibe+1)
    real   (rp), target, intent(in)    :: k      (ips-2:ipe+2)
    real   (rp), intent(in)    :: dt
    real   (rp), intent(in)    :: bvalue (2)
    real   (rp), intent(inout) :: flux   (2)
    real(c_double) :: bv(2), fl(2)
    integer(c_int) :: bc(2), rc
    if (.not. c_associated(CB%ctx)) call ensure_ctx(1_ip)
    bv = real(bvalue, c_double)
    bc = int(bcode, c_int)
    rc = f3d_solve_1d(CB%ctx, c_loc(c), c_loc(dxp), c_loc(dxb), c_loc(u), c_loc(k), real(dt, c_double), &
         bv, fl, bc, int(ips, c_int), int(ipe, c_int), int(limiter, c_int))
    call check(rc, 'fall3d_ads_solve_1d')
    flux = real(fl, rp)
  end subroutine ADS_solve_1D
  !
END MODULE ADS
