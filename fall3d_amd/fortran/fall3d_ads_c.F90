!***************************************************************
!> fall3d_ads_c: ISO_C_BINDING interfaces of the resident group of the
!> C ABI (include/fall3d_ads.h), for Fortran host code that keeps the
!> concentration, meteo and deposit on the GPU across time steps: the time
!> loop of Sources/mod_F3D.F90 (F3D_update_meteo_term :564-969,
!> F3D_set_source_term, F3D_time_step :1397-1491, F3D_end_time_step
!> :1235-1388), the output of nc_IO_out_res (Sources/mod_nc_IO.F90:3206-4335),
!> restart snapshots and the ensemble members of task_SetEns.
!>
!> Host arrays are passed with c_loc() in the layouts and kind (real_kind)
!> the header documents; the derived types are bind(C) copies of the header's
!> argument structs (same member order). Every function returns 0 or an error
!> code; fall3d_ads_error(ctx) returns the message as a Fortran string.
!> The drop-in MODULE ADS (mod_ADS_hip.F90) is independent of this module.
!***************************************************************
module fall3d_ads_c
  use iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: FALL3D_ADS_STATE_FP32 = 0, FALL3D_ADS_STATE_FP32X2 = 1, &
                                       FALL3D_ADS_STATE_FP64 = 2
  integer(c_int), parameter, public :: FALL3D_ADS_BC_FREEFLOW = 0, FALL3D_ADS_BC_DIRICHLET = 1, &
                                       FALL3D_ADS_BC_PERIODIC = 2
  integer, parameter, public :: FALL3D_ADS_BUDGET_LEN = 12

  !> meteo update arguments (fall3d_meteo_options / _grid / _corner)
  type, bind(C), public :: fall3d_meteo_options
     integer(c_int) :: map_h, map_v, modkh, modkv, modv, nbins
     real(c_double) :: kh0, kv0, rams_cs, dlon, dlat, x3max, s_time, vset_fac
     type(c_ptr)    :: diam = c_null_ptr, rho = c_null_ptr, psi = c_null_ptr
     type(c_ptr)    :: cat = c_null_ptr, aggregate = c_null_ptr
     integer(c_int) :: dry_deposition = 0
  end type fall3d_meteo_options

  type, bind(C), public :: fall3d_meteo_grid
     type(c_ptr) :: Hm1_p = c_null_ptr, Hm2_p = c_null_ptr, Hm3_p = c_null_ptr
     type(c_ptr) :: dhdx_p = c_null_ptr, dhdy_p = c_null_ptr, z_c = c_null_ptr, h_c = c_null_ptr
     type(c_ptr) :: X3_c = c_null_ptr, X3_p = c_null_ptr
  end type fall3d_meteo_grid

  type, bind(C), public :: fall3d_meteo_corner
     type(c_ptr) :: uc(2) = c_null_ptr, vc(2) = c_null_ptr, wc(2) = c_null_ptr
     type(c_ptr) :: tc(2) = c_null_ptr, tvc(2) = c_null_ptr, rhoc(2) = c_null_ptr
     type(c_ptr) :: pblhc(2) = c_null_ptr, ustc(2) = c_null_ptr, monc(2) = c_null_ptr
     type(c_ptr) :: lusec = c_null_ptr
  end type fall3d_meteo_corner

  !> output post-processing arguments (fall3d_output_args, fall3d_cut_args)
  type, bind(C), public :: fall3d_output_args
     type(c_ptr)    :: Hm1_p = c_null_ptr, Hm2_p = c_null_ptr, Hm3_p = c_null_ptr, z_c = c_null_ptr
     type(c_ptr)    :: bin_species = c_null_ptr
     integer(c_int) :: species = 0, species_category = 1, so2_du = 0, npm = 0
     type(c_ptr)    :: pm_um = c_null_ptr, bin_diam = c_null_ptr
  end type fall3d_output_args

  type, bind(C), public :: fall3d_cut_args
     real(c_double) :: lonmin = 0.0_c_double, dlon = 1.0_c_double, latmin = 0.0_c_double, dlat = 1.0_c_double
     type(c_ptr)    :: sigma = c_null_ptr
     real(c_double) :: X3max = 0.0_c_double
     integer(c_int) :: ncutx = 0, ncuty = 0, ncutz = 0, nfl = 0
     type(c_ptr)    :: x_cut = c_null_ptr, y_cut = c_null_ptr, z_cut = c_null_ptr, fl_cut = c_null_ptr
  end type fall3d_cut_args

  public :: fall3d_ads_error

  interface
     ! ---- context -------------------------------------------------------
     integer(c_int) function fall3d_ads_initialize(ctx, nx, ny, nz, nbins, limiter, time_marching, real_kind, &
                                                   state_mode) bind(C, name="fall3d_ads_initialize")
       import :: c_ptr, c_int
       type(c_ptr), intent(out) :: ctx
       integer(c_int), value :: nx, ny, nz, nbins, limiter, time_marching, real_kind, state_mode
     end function fall3d_ads_initialize
     integer(c_int) function fall3d_ads_release(ctx) bind(C, name="fall3d_ads_release")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
     end function fall3d_ads_release
     type(c_ptr) function fall3d_ads_last_error(ctx) bind(C, name="fall3d_ads_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function fall3d_ads_last_error
     integer(c_int) function fall3d_ads_set_grid(ctx, dX1_p, dX1_b, dX2_p, dX2_b, dX3_p, dX3_b) &
          bind(C, name="fall3d_ads_set_grid")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, dX1_p, dX1_b, dX2_p, dX2_b, dX3_p, dX3_b
     end function fall3d_ads_set_grid
     integer(c_int) function fall3d_ads_set_boundary(ctx, mode) bind(C, name="fall3d_ads_set_boundary")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
       integer(c_int), value :: mode
     end function fall3d_ads_set_boundary
     integer(c_int) function fall3d_ads_synchronize(ctx) bind(C, name="fall3d_ads_synchronize")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx
     end function fall3d_ads_synchronize

     ! ---- meteo ---------------------------------------------------------
     integer(c_int) function fall3d_ads_set_meteo(ctx, u1, u2, v1, v2, w1, w2, k1, k2, k3, vs) &
          bind(C, name="fall3d_ads_set_meteo")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, u1, u2, v1, v2, w1, w2, k1, k2, k3, vs
     end function fall3d_ads_set_meteo
     integer(c_int) function fall3d_ads_update_meteo(ctx, corner, grid, opts) bind(C, name="fall3d_ads_update_meteo")
       import :: c_ptr, c_int, fall3d_meteo_corner, fall3d_meteo_grid, fall3d_meteo_options
       type(c_ptr), value :: ctx
       type(fall3d_meteo_corner), intent(in) :: corner
       type(fall3d_meteo_grid), intent(in) :: grid
       type(fall3d_meteo_options), intent(in) :: opts
     end function fall3d_ads_update_meteo
     integer(c_int) function fall3d_ads_get_meteo(ctx, u1, u2, v1, v2, w1, w2, k1, k2, k3, vs) &
          bind(C, name="fall3d_ads_get_meteo")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, u1, u2, v1, v2, w1, w2, k1, k2, k3, vs
     end function fall3d_ads_get_meteo
     integer(c_int) function fall3d_ads_get_time_step(ctx, stime, criterion, cfl, dt) &
          bind(C, name="fall3d_ads_get_time_step")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: stime, cfl
       integer(c_int), value :: criterion
       real(c_double), intent(out) :: dt
     end function fall3d_ads_get_time_step

     ! ---- source, state, stepping ---------------------------------------
     integer(c_int) function fall3d_ads_set_source(ctx, s) bind(C, name="fall3d_ads_set_source")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, s
     end function fall3d_ads_set_source
     integer(c_int) function fall3d_ads_set_source_metric(ctx, s, Hm1_p, Hm2_p, Hm3_p) &
          bind(C, name="fall3d_ads_set_source_metric")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, s, Hm1_p, Hm2_p, Hm3_p
     end function fall3d_ads_set_source_metric
     integer(c_int) function fall3d_ads_set_state(ctx, c, acum) bind(C, name="fall3d_ads_set_state")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, c, acum
     end function fall3d_ads_set_state
     integer(c_int) function fall3d_ads_get_state(ctx, c, acum) bind(C, name="fall3d_ads_get_state")
       import :: c_ptr, c_int
       type(c_ptr), value :: ctx, c, acum
     end function fall3d_ads_get_state
     integer(c_int) function fall3d_ads_step(ctx, dt, stime, iiter) bind(C, name="fall3d_ads_step")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt, stime
       integer(c_int), value :: iiter
     end function fall3d_ads_step
     integer(c_int) function fall3d_ads_run(ctx, dt, stime, iiter0, nsteps) bind(C, name="fall3d_ads_run")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt, stime
       integer(c_int), value :: iiter0, nsteps
     end function fall3d_ads_run
     integer(c_int) function fall3d_ads_get_fluxes(ctx, flux) bind(C, name="fall3d_ads_get_fluxes")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: flux(6)
     end function fall3d_ads_get_fluxes
     integer(c_int) function fall3d_ads_set_fluxes(ctx, flux) bind(C, name="fall3d_ads_set_fluxes")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: flux(6)
     end function fall3d_ads_set_fluxes
     integer(c_int) function fall3d_ads_get_budget(ctx, budget) bind(C, name="fall3d_ads_get_budget")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: budget(12)
     end function fall3d_ads_get_budget
     integer(c_int) function fall3d_ads_check_state(ctx, counts, min_value) bind(C, name="fall3d_ads_check_state")
       import :: c_ptr, c_int, c_double, c_long_long
       type(c_ptr), value :: ctx
       integer(c_long_long), intent(out) :: counts(2)
       real(c_double), intent(out) :: min_value
     end function fall3d_ads_check_state
     integer(c_int) function fall3d_ads_autotune(ctx, dt, stime, reps, tiling, best_ms) &
          bind(C, name="fall3d_ads_autotune")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt, stime
       integer(c_int), value :: reps
       integer(c_int), intent(out) :: tiling(3)
       real(c_double), intent(out) :: best_ms(3)
     end function fall3d_ads_autotune

     ! ---- output and restart --------------------------------------------
     integer(c_int) function fall3d_ads_postprocess(ctx, args, con_bins, con, con_sur, col_mass, col_mass_pm, &
                                                    cloud_top, grn_load, grn_load_bin) &
          bind(C, name="fall3d_ads_postprocess")
       import :: c_ptr, c_int, fall3d_output_args
       type(c_ptr), value :: ctx
       type(fall3d_output_args), intent(in) :: args
       type(c_ptr), value :: con_bins, con, con_sur, col_mass, col_mass_pm, cloud_top, grn_load, grn_load_bin
     end function fall3d_ads_postprocess
     integer(c_int) function fall3d_ads_postprocess_cuts(ctx, args, cuts, cut_x, cut_y, cut_z, cut_fl) &
          bind(C, name="fall3d_ads_postprocess_cuts")
       import :: c_ptr, c_int, fall3d_output_args, fall3d_cut_args
       type(c_ptr), value :: ctx
       type(fall3d_output_args), intent(in) :: args
       type(fall3d_cut_args), intent(in) :: cuts
       type(c_ptr), value :: cut_x, cut_y, cut_z, cut_fl
     end function fall3d_ads_postprocess_cuts
     integer(c_int) function fall3d_ads_snapshot(ctx, buf, bytes, time, iiter) bind(C, name="fall3d_ads_snapshot")
       import :: c_ptr, c_int, c_size_t, c_double, c_long_long
       type(c_ptr), value :: ctx, buf
       integer(c_size_t), intent(inout) :: bytes
       real(c_double), value :: time
       integer(c_long_long), value :: iiter
     end function fall3d_ads_snapshot
     integer(c_int) function fall3d_ads_restore(ctx, buf, bytes, time, iiter) bind(C, name="fall3d_ads_restore")
       import :: c_ptr, c_int, c_size_t, c_double, c_long_long
       type(c_ptr), value :: ctx, buf
       integer(c_size_t), value :: bytes
       real(c_double), intent(out) :: time
       integer(c_long_long), intent(out) :: iiter
     end function fall3d_ads_restore

     ! ---- ensembles -----------------------------------------------------
     integer(c_int) function fall3d_ads_ensemble_create(ens, nmembers, dims, nbins, limiter, time_marching, &
                                                        real_kind, state_mode) bind(C, name="fall3d_ads_ensemble_create")
       import :: c_ptr, c_int
       type(c_ptr), intent(out) :: ens
       integer(c_int), value :: nmembers, nbins, limiter, time_marching, real_kind, state_mode
       integer(c_int), intent(in) :: dims(*)
     end function fall3d_ads_ensemble_create
     integer(c_int) function fall3d_ads_ensemble_member(ens, m, ctx) bind(C, name="fall3d_ads_ensemble_member")
       import :: c_ptr, c_int
       type(c_ptr), value :: ens
       integer(c_int), value :: m
       type(c_ptr), intent(out) :: ctx
     end function fall3d_ads_ensemble_member
     integer(c_int) function fall3d_ads_ensemble_step(ens, dt, stime, iiter) bind(C, name="fall3d_ads_ensemble_step")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ens
       real(c_double), intent(in) :: dt(*), stime(*)
       integer(c_int), value :: iiter
     end function fall3d_ads_ensemble_step
     integer(c_int) function fall3d_ads_ensemble_run(ens, dt, stime, iiter0, nsteps) &
          bind(C, name="fall3d_ads_ensemble_run")
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: ens
       real(c_double), intent(in) :: dt(*), stime(*)
       integer(c_int), value :: iiter0, nsteps
     end function fall3d_ads_ensemble_run
     integer(c_int) function fall3d_ads_ensemble_synchronize(ens) bind(C, name="fall3d_ads_ensemble_synchronize")
       import :: c_ptr, c_int
       type(c_ptr), value :: ens
     end function fall3d_ads_ensemble_synchronize
     integer(c_int) function fall3d_ads_ensemble_release(ens) bind(C, name="fall3d_ads_ensemble_release")
       import :: c_ptr, c_int
       type(c_ptr), value :: ens
     end function fall3d_ads_ensemble_release
  end interface

  public :: fall3d_ads_initialize, fall3d_ads_release, fall3d_ads_last_error, fall3d_ads_set_grid, &
            fall3d_ads_set_boundary, fall3d_ads_synchronize, fall3d_ads_set_meteo, fall3d_ads_update_meteo, &
            fall3d_ads_get_meteo, fall3d_ads_get_time_step, fall3d_ads_set_source, fall3d_ads_set_source_metric, &
            fall3d_ads_set_state, fall3d_ads_get_state, fall3d_ads_step, fall3d_ads_run, fall3d_ads_get_fluxes, &
            fall3d_ads_set_fluxes, fall3d_ads_get_budget, fall3d_ads_check_state, fall3d_ads_autotune, &
            fall3d_ads_postprocess, fall3d_ads_postprocess_cuts, fall3d_ads_snapshot, fall3d_ads_restore, &
            fall3d_ads_ensemble_create, fall3d_ads_ensemble_member, fall3d_ads_ensemble_step, &
            fall3d_ads_ensemble_run, fall3d_ads_ensemble_synchronize, fall3d_ads_ensemble_release

contains

  !> the context's last error message as a Fortran string
  function fall3d_ads_error(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: n
    p = fall3d_ads_last_error(ctx)
    if (.not. c_associated(p)) then
       msg = ''
       return
    end if
    call c_f_pointer(p, s, [4096])
    n = 0
    do while (n < 4096)
       if (s(n + 1) == c_null_char) exit
       n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function fall3d_ads_error

end module fall3d_ads_c
