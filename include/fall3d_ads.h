/*
 * fall3d_ads.h -- C ABI of the MI355X (gfx950) ADS step library
 * (libfall3d_ads_hip.so).
 *
 * The library replaces FALL3D's MODULE ADS (Sources/mod_ADS.F90) and the
 * ADS part of the time step around it. Two groups of entry points:
 *
 *  1. Drop-in group: one C function per public procedure of MODULE ADS
 *     (Sources/mod_ADS.F90:48-55), same argument meaning and order, host
 *     arrays in the reference's layouts and kind (real(rp), rp = 4 or 8).
 *     Each call copies its arrays to the GPU, runs the sweep there and
 *     copies the results back (correct, but PCIe-bound). This is what
 *     fall3d_amd/fortran/mod_ADS_hip.F90 binds through ISO_C_BINDING so
 *     that Sources/mod_F3D.F90 compiles and runs unchanged.
 *
 *  2. Resident group (no counterpart procedure in the reference; it fuses
 *     what Sources/mod_F3D.F90:1397-1491 does around the ADS calls): the
 *     device owns the concentration field, the settling velocities and the
 *     wind/diffusion snapshots; a whole F3D_time_step (source injection,
 *     x/y sweeps in iiter-parity order, z sweep with ground deposition) is
 *     enqueued by fall3d_ads_step without host traffic. Host copies happen
 *     only at fall3d_ads_set_state / get_state / get_budget.
 *
 * Array layouts (Fortran column-major; bounds as in Sources/mod_Domain.F90
 * for the whole single-GPU domain, ips=1, ipe=nx, 2 ghost cells, 1 halo face):
 *   c    (-1:nx+2,-1:ny+2,-1:nz+2,nbins)   k1 (-1:nx+2,1:ny,1:nz)
 *   u    ( 0:nx+2, 1:ny,   1:nz)           k2 (-1:ny+2,1:nx,1:nz)
 *   v    ( 0:ny+2, 1:nx,   1:nz)           k3 (-1:nz+2,1:nx,1:ny)
 *   w    ( 0:nz+2, 1:nx,   1:ny)           vs ( 1:nx,1:ny,0:nz+2,nbins)
 *   acum (-1:nx+2,-1:ny+2,nbins)           s  ( 1:nx,1:ny,1:nz,nbins)
 *   dX1_p(-1:nx+2) dX1_b(0:nx+2), likewise dX2_*, dX3_*
 * Every array pointer is a HOST pointer of kind real_kind (4 or 8 bytes)
 * unless stated otherwise; flux accumulators are always double.
 *
 * All functions return FALL3D_ADS_OK (0) or a positive error code; the
 * message is available from fall3d_ads_last_error. A context is not
 * thread-safe (one host thread per context), as MODULE ADS is a singleton.
 */
#ifndef FALL3D_ADS_H
#define FALL3D_ADS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fall3d_ads_ctx fall3d_ads_ctx;

/* limiter (Sources/mod_ADS.F90:121-131) */
#define FALL3D_ADS_MINMOD 1
#define FALL3D_ADS_SUPERBEE 2
#define FALL3D_ADS_OSPRE 3
/* time marching (Sources/mod_ADS.F90:43-44) */
#define FALL3D_ADS_EULER 1
#define FALL3D_ADS_RK 2
/* device state of the concentration: plain FP32, or FP32 hi + FP32 lo
 * (compensated; telescoping flux updates are then exact to ~2^-48 and the
 * mass budget closes to FP64 rounding), or FP64 with FP64 fields: the
 * reference's arithmetic statement by statement in double precision, for
 * truth runs and real(8) drop-in use. The FP64 state supports the drop-in
 * sweeps, set/get state, source, step/run (graphs included), time step,
 * fluxes, budget, snapshot/restore and check_state; the tile controls, the
 * allocation-free step, face-flux output and the learned correction are
 * FP32-state features and return EINVAL for it. */
#define FALL3D_ADS_STATE_FP32 0
#define FALL3D_ADS_STATE_FP32X2 1
#define FALL3D_ADS_STATE_FP64 2

/* status codes */
#define FALL3D_ADS_OK 0
#define FALL3D_ADS_EINVAL 1
#define FALL3D_ADS_EHIP 2
#define FALL3D_ADS_ESTATE 3

/* budget vector of fall3d_ads_get_budget */
#define FALL3D_ADS_BUDGET_MASS_VOLUME 0 /* sum c dX1_p dX2_p dX3_p (Sources/mod_Grid.F90:1364-1375) */
#define FALL3D_ADS_BUDGET_MASS_CELLS 1  /* sum c dX1_b dX2_b dX3_b (measure conserved by the flux form) */
#define FALL3D_ADS_BUDGET_GROUND 2      /* |D_flux| (Sources/mod_Grid.F90:1488) */
#define FALL3D_ADS_BUDGET_LATERAL 3     /* |W|+|E|+|S|+|N|+|U| (Sources/mod_Grid.F90:1499-1500) */
#define FALL3D_ADS_BUDGET_INJECTED 4    /* mass added by source injection */
#define FALL3D_ADS_BUDGET_LEDGER_W 5    /* mass that left through W,E,S,N,D,U (RK-weighted, exact) */
/* FP32 state: the mass that rounding the sweeps' final updates to FP32 took
 * from the stored field (sum of (FP64 update - stored FP32 value) dxb area);
 * MASS_CELLS + ledger + RESIDUAL - INJECTED stays at its initial value to
 * FP64 rounding. 0 with the compensated state. */
#define FALL3D_ADS_BUDGET_RESIDUAL 11
#define FALL3D_ADS_BUDGET_LEN 12

/* ---------------------------------------------------------------------
 * Drop-in group (MODULE ADS)
 * --------------------------------------------------------------------- */

/* ADS_initialize(limitertype, time_marching), Sources/mod_ADS.F90:98-157.
 * The grid extents replace the Domain module globals the reference reads
 * (my_ips..my_kpe, Sources/mod_Domain.F90:263-330); real_kind is the host
 * real(rp) kind (4 or 8). state_mode selects the device state
 * (FALL3D_ADS_STATE_*). */
int fall3d_ads_initialize(fall3d_ads_ctx **ctx, int nx, int ny, int nz, int nbins, int limiter,
                          int time_marching, int real_kind, int state_mode);
/* Size limit: one bin of c with its ghosts, (nx+4)(ny+4)(nz+4) floats, and
 * each wind/diffusion field must stay below 2 GiB (the sweeps address them
 * with 32-bit buffer offsets); larger grids are rejected with EINVAL. */

/* The ARAKAWA_C_GRID spacing arrays ADS reads (Sources/mod_KindType.F90:486-491):
 * dX1_p(-1:nx+2), dX1_b(0:nx+2), ... Must be called before any solve. */
int fall3d_ads_set_grid(fall3d_ads_ctx *ctx, const void *dX1_p, const void *dX1_b, const void *dX2_p,
                        const void *dX2_b, const void *dX3_p, const void *dX3_b);

/* ADS_get_vcomps(my_u, my_v, my_w), Sources/mod_ADS.F90:166-177: returns
 * the host arrays CB%u (0:nx+2,1:ny,1:nz), CB%v, CB%w that the caller fills
 * before the sweeps (Sources/mod_F3D.F90:1435-1439). They are owned by the
 * context and uploaded at the next solve. */
int fall3d_ads_get_vcomps(fall3d_ads_ctx *ctx, void **u, void **v, void **w);

/* ADS_update_horizontal_vcomps(u1,u2,v1,v2,stime), Sources/mod_ADS.F90:187-198:
 * CB%u = (1-stime) u1 + stime u2, CB%v likewise (computed on the GPU). */
int fall3d_ads_update_horizontal_vcomps(fall3d_ads_ctx *ctx, const void *u1, const void *u2, const void *v1,
                                        const void *v2, double stime);

/* ADS_solve_along_x(W_flux,E_flux,dt,c,k1,GRID,nbins), Sources/mod_ADS.F90:233-412.
 * c is updated in place (interior and the free-flow x ghosts); W/E_flux are
 * incremented by the boundary mass fluxes of the last RK stage. */
int fall3d_ads_solve_along_x(fall3d_ads_ctx *ctx, double *W_flux, double *E_flux, double dt, void *c,
                             const void *k1);

/* ADS_solve_along_y(S_flux,N_flux,dt,c,k2,GRID,nbins), Sources/mod_ADS.F90:422-602. */
int fall3d_ads_solve_along_y(fall3d_ads_ctx *ctx, double *S_flux, double *N_flux, double dt, void *c,
                             const void *k2);

/* ADS_solve_along_z(D_flux,U_flux,dt,c,acum,k3,w1,w2,stime,vs,GRID,nbins),
 * Sources/mod_ADS.F90:612-811. CB%w is formed as (1-stime) w1 + stime w2 - vs
 * per bin; acum += |ground flux| dt. */
int fall3d_ads_solve_along_z(fall3d_ads_ctx *ctx, double *D_flux, double *U_flux, double dt, void *c,
                             void *acum, const void *k3, const void *w1, const void *w2, double stime,
                             const void *vs);

/* ADS_solve_1D(c,dxp,dxb,u,k,dt,bvalue,flux,bcode,ips,ipe,ibs,ibe,limiter,idime),
 * Sources/mod_ADS.F90:820-1053: one line integrated with RK4 and per-side
 * boundary code bcode(2) in {0 none, -1 periodic, 1 Dirichlet, 2 free flow}
 * (idime is accepted and ignored, as in the reference). Arrays:
 * c(ips-2:ipe+2), dxp(ips-2:ipe+2), dxb(ibs-1:ibe+1), u(ibs-1:ibe+1),
 * k(ips-2:ipe+2) of real_kind; flux(2) double. Runs on the GPU. */
int fall3d_ads_solve_1d(fall3d_ads_ctx *ctx, void *c, const void *dxp, const void *dxb, const void *u,
                        const void *k, double dt, const double bvalue[2], double flux[2], const int bcode[2],
                        int ips, int ipe, int limiter);

/* ADS_release(), Sources/mod_ADS.F90:208-224. Frees every device resource. */
int fall3d_ads_release(fall3d_ads_ctx *ctx);

/* Message of the last failing call on this context (never NULL). */
const char *fall3d_ads_last_error(const fall3d_ads_ctx *ctx);

/* ---------------------------------------------------------------------
 * Resident group (fuses Sources/mod_F3D.F90:1397-1491 around the sweeps)
 * --------------------------------------------------------------------- */

/* Upload one meteo interval: the two snapshots of the scaled face winds
 * MY_MET%my_u/v/w(...,1:2) and the diffusion/settling fields
 * (Sources/mod_F3D.F90:916-944). Any of u2/v2/w2 may be NULL (= snapshot 1). */
int fall3d_ads_set_meteo(fall3d_ads_ctx *ctx, const void *u1, const void *u2, const void *v1, const void *v2,
                         const void *w1, const void *w2, const void *k1, const void *k2, const void *k3,
                         const void *vs);

/* Boundary rule of the sweeps' ghost cells on every physical side.
 * ADS_initialize hard-wires free flow (Sources/mod_ADS.F90:116); the
 * reference also holds dirichlet (:1330-1399: ghost = 2 bvalue - c, bvalue
 * 0) and periodic (:1403-1448: in the reference a zero-gradient copy of the
 * first cell, not a wrap-around) behind the same procedure pointer. The
 * outer ghost copies the inner one in every rule. Default FREEFLOW. */
#define FALL3D_ADS_BC_FREEFLOW 0
#define FALL3D_ADS_BC_DIRICHLET 1
#define FALL3D_ADS_BC_PERIODIC 2
int fall3d_ads_set_boundary(fall3d_ads_ctx *ctx, int mode);

/* ---- 2a. Meteo update on the device (F3D_update_meteo_term,
 * Sources/mod_F3D.F90:564-969) --------------------------------------------
 * From the meteo fields at the cell corners of the two meteo times that
 * bound the interval (MY_MET%my_uc(...,its/ite) and friends) to the fields
 * set_meteo takes, computed on the device and left there: ground corner
 * winds zeroed (:662-669), corner -> face winds with one ghost face
 * (Sources/mod_Grid.F90:685-752), their metric scaling (:761-870), the
 * horizontal and vertical eddy diffusivities (Sources/mod_Phys.F90:406-653),
 * the settling velocities of every bin (:673-960) and the scaling of k1, k2,
 * k3 and v_s (Sources/mod_Grid.F90:880-1017). Diffusion and settling use the
 * corner fields interpolated to s_time, (1 - s_time) its + s_time ite, as
 * the reference does at the middle of the interval (mod_F3D.F90:704-705).
 * Every value is formed in FP64; FP32 states store it rounded. Dry
 * deposition velocities (phys_dry_deposition) are not added. Codes follow
 * the reference's parameters. */
typedef struct fall3d_meteo_options {
  int map_h;    /* 0 Cartesian, 1 spherical (Sources/mod_Grid.F90:17-18) */
  int map_v;    /* 0 Cartesian, 1 sigma without decay, 2 sigma with linear decay (:22-24) */
  int modkh;    /* horizontal diffusion: 1 constant, 2 CMAQ, 4 RAMS (Sources/mod_Phys.F90:53-56) */
  int modkv;    /* vertical diffusion: 1 constant, 3 similarity */
  int modv;     /* settling model 1..7: Arastopour, Ganser, Wilson, Dellino, Pfeiffer,
                 * Dioguardi 2017, Dioguardi 2018 (:44-50) */
  int nbins;    /* must equal the context's */
  double kh0, kv0;    /* constant diffusivities (MY_MOD%kh0, kv0) */
  double rams_cs;     /* MY_MOD%RAMS_CS */
  double dlon, dlat;  /* MY_GRID%dlon, dlat (CMAQ) */
  double x3max;       /* MY_GRID%X3max (sigma with linear decay) */
  double s_time;      /* interpolation factor of the diffusion and settling inputs */
  double vset_fac;    /* MY_AGR%vset_fac, applied to aggregate bins */
  const double *diam, *rho, *psi;  /* per bin: diameter, density, shape factor (MY_BIN) */
  const int *cat;                  /* per bin: category (2 = aerosol: v_s = 0) */
  const int *aggregate;            /* per bin: 1 for an aggregate bin */
  int dry_deposition;              /* 1: add the dry deposition velocity (Feng 2008,
                                    * Sources/mod_Phys.F90:1479-1673) to v_s at the ground
                                    * face and the one below it, before scaling
                                    * (Sources/mod_F3D.F90:790-801); needs ustc, lusec, z_c */
} fall3d_meteo_options;

/* Grid arrays of the update (host arrays of real_kind). */
typedef struct fall3d_meteo_grid {
  const void *Hm1_p;           /* (1:ny) */
  const void *Hm2_p;           /* (-1:ny+2) */
  const void *Hm3_p;           /* (1:nx,1:ny) */
  const void *dhdx_p, *dhdy_p; /* (1:nx,1:ny); read for map_v 1, 2 */
  const void *z_c;             /* (1:nx+1,1:ny+1,1:nz+1) corner heights; read for similarity */
  const void *h_c;             /* (1:nx+1,1:ny+1) corner topography; read for similarity */
  const void *X3_c;            /* (1:nz+1); read for map_v 2 */
  const void *X3_p;            /* (1:nz); read for map_v 2 */
} fall3d_meteo_grid;

/* Corner fields (host arrays of real_kind), [0] at the interval's first
 * meteo time, [1] at its second: 3-D (1:nx+1,1:ny+1,1:nz+1), 2-D
 * (1:nx+1,1:ny+1). tvc, pblhc, ustc, monc are read for similarity only. */
typedef struct fall3d_meteo_corner {
  const void *uc[2], *vc[2], *wc[2];
  const void *tc[2], *tvc[2], *rhoc[2];
  const void *pblhc[2], *ustc[2], *monc[2];
  const void *lusec; /* (1:nx+1,1:ny+1) USGS land use (MY_MET%my_lusec); read with dry_deposition */
} fall3d_meteo_corner;

/* Runs the update and makes its output the context's meteo (as set_meteo);
 * returns ESTATE when a settling velocity does not converge (as the
 * reference stops with 'no convergence in vset'). */
int fall3d_ads_update_meteo(fall3d_ads_ctx *ctx, const fall3d_meteo_corner *corner, const fall3d_meteo_grid *grid,
                            const fall3d_meteo_options *opts);
/* Reads the context's meteo back in the layouts set_meteo takes (any
 * pointer may be NULL): u1/u2 (nx+3,ny,nz), v1/v2 (ny+3,nx,nz), w1/w2
 * (nz+3,nx,ny), k1 (nx+4,ny,nz), k2 (ny+4,nx,nz), k3 (nz+4,nx,ny),
 * vs (nx,ny,nz+3,nbins), real_kind. */
int fall3d_ads_get_meteo(fall3d_ads_ctx *ctx, void *u1, void *u2, void *v1, void *v2, void *w1, void *w2, void *k1,
                         void *k2, void *k3, void *vs);

/* ---- 2b. Output post-processing (nc_IO_out_res, Sources/mod_nc_IO.F90:
 * 3206-4335) -----------------------------------------------------------------
 * From the context's state, on the device: c scaled back by the map
 * factors, converted to g/m3 and clipped at 0 (:3397-3416), averaged to the
 * cell corners (Sources/mod_Grid.F90:1150-1181, ghosts included); the ground
 * accumulation averaged to the corners (:1190-1220, kg/m2); and for one
 * species the reference's result variables: CON, CON_SUR, COL_MASS (g/m2, or
 * Dobson units with so2_du: / 64 x 2.238e3), COL_MASS_PM (bins with diameter
 * <= each threshold; none for an aerosol species), CLOUD_TOP (height of the
 * highest corner above 50 ug/m3, else 0) and GRN_LOAD, plus CON_BIN and
 * GRN_LOAD_BIN for every bin (mod_nc_IO.F90:3506-4020). Wet deposition is
 * not part of the ground load here (no wet deposition runs on the device);
 * the z and flight-level cuts, which interpolate these fields for the
 * output files, are left to the caller. */
typedef struct fall3d_output_args {
  const void *Hm1_p, *Hm2_p, *Hm3_p; /* (1:ny), (-1:ny+2), (1:nx,1:ny) of real_kind; NULL = 1 */
  const void *z_c;                   /* corner heights (1:nx+1,1:ny+1,1:nz+1), real_kind */
  const int *bin_species;            /* species code of each bin; NULL = every bin one species */
  int species;                       /* the species the species products sum */
  int species_category;              /* its category (2 = aerosol: no PM loads) */
  int so2_du;                        /* 1: COL_MASS in Dobson units */
  int npm;                           /* number of PM thresholds (at most 8) */
  const double *pm_um;               /* PM thresholds in micrometres */
  const double *bin_diam;            /* diameter of each bin (m) */
} fall3d_output_args;

/* Outputs are FP64 host arrays at the corners, x fastest (any may be NULL):
 * con_bins (nx+1,ny+1,nz+1,nbins), con (nx+1,ny+1,nz+1), con_sur, col_mass,
 * cloud_top, grn_load (nx+1,ny+1), col_mass_pm (nx+1,ny+1,npm), grn_load_bin
 * (nx+1,ny+1,nbins). */
int fall3d_ads_postprocess(fall3d_ads_ctx *ctx, const fall3d_output_args *args, double *con_bins, double *con,
                           double *con_sur, double *col_mass, double *col_mass_pm, double *cloud_top,
                           double *grn_load, double *grn_load_bin);

/* Concentration cuts of nc_IO_out_res (Sources/mod_nc_IO.F90:3292-3393
 * and 4084-4332) for the species of args: the species' corner
 * concentration (CON, g/m3) is interpolated per corner column from the
 * column's heights z_c onto the levels sigma(k) X3max, then linearly to the
 * cut positions; each position's interval and weight are those of
 * grid_get_shapez (Sources/mod_Grid.F90:1287-1329), so a position outside
 * the mesh takes the end value (the reference then only warns). */
typedef struct fall3d_cut_args {
  double lonmin, dlon;   /* corner i (1..nx+1) at lonmin + (i-1) dlon (MY_GRID) */
  double latmin, dlat;   /* corner j (1..ny+1) at latmin + (j-1) dlat */
  const double *sigma;   /* gl_sigma(1:nz+1): the corner sigma levels */
  double X3max;          /* domain top (m) */
  int ncutx, ncuty, ncutz, nfl;
  const double *x_cut, *y_cut, *z_cut; /* cut positions (x, y in the grid's units; z in m) */
  const double *fl_cut;                /* flight levels (hundreds of feet: z = 30.48 FL m) */
} fall3d_cut_args;

/* Outputs (FP64 host arrays, a count of 0 leaves its array untouched and it
 * may be NULL): cut_x (ny+1,nz+1,ncutx), cut_y (nx+1,nz+1,ncuty), cut_z
 * (nx+1,ny+1,ncutz), cut_fl (nx+1,ny+1,nfl), first index fastest. */
int fall3d_ads_postprocess_cuts(fall3d_ads_ctx *ctx, const fall3d_output_args *args, const fall3d_cut_args *cuts,
                                double *cut_x, double *cut_y, double *cut_z, double *cut_fl);

/* Dense source rate s(1:nx,1:ny,1:nz,nbins) (MY_TRA%my_s), injected before
 * the sweeps as c += dt s / (dX1_p dX2_p dX3_p Hm1 Hm2 Hm3)
 * (Sources/mod_F3D.F90:1456-1473); stored sparsely on the device (only
 * nonzero entries are injected). NULL clears it. set_source takes all map
 * factors as 1: that is FALL3D's flat Cartesian grid (Sources/mod_Grid.F90:
 * 403,463,603). On spherical or terrain-following grids pass the grid's
 * MY_GRID%Hm1_p(1:ny), Hm2_p(-1:ny+2) and Hm3_p(1:nx,1:ny) (real_kind, any
 * may be NULL = 1) to set_source_metric. */
int fall3d_ads_set_source(fall3d_ads_ctx *ctx, const void *s);
int fall3d_ads_set_source_metric(fall3d_ads_ctx *ctx, const void *s, const void *Hm1_p, const void *Hm2_p,
                                 const void *Hm3_p);

/* Concentration (with ghosts) and ground accumulation; NULL sets zero. */
int fall3d_ads_set_state(fall3d_ads_ctx *ctx, const void *c, const void *acum);
int fall3d_ads_get_state(fall3d_ads_ctx *ctx, void *c, void *acum);

/* grid_get_time_step (Sources/mod_Grid.F90:1026-1141) on the GPU with an
 * FP64 max-reduction, for the winds interpolated at stime. criterion 1 or 2. */
int fall3d_ads_get_time_step(fall3d_ads_ctx *ctx, double stime, int criterion, double cfl, double *dt);

/* Enqueue one F3D_time_step (source add, sweeps x/y ordered by iiter parity,
 * z) on the context stream. Asynchronous. */
int fall3d_ads_step(fall3d_ads_ctx *ctx, double dt, double stime, int iiter);

/* fall3d_ads_step without any device allocation. No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The per-workgroup partial sums of the step's source and sweep kernels go
 * to a static device array of 2^21 doubles (16 MiB) per GPU, held until the
 * process ends, instead of the context's buffers; the kernels write every
 * partial they later reduce, so nothing needs clearing. A step that needs
 * more partial rows than that is rejected with EINVAL (the message gives
 * both sizes); so is a context with a learned correction set or profiling
 * on. Every call on one GPU shares that GPU's array: calls that overlap on
 * one GPU (two contexts on different streams, or the same context from two
 * threads) race and must be serialized by the caller. Results are bitwise
 * those of fall3d_ads_step. */
int fall3d_ads_step_noalloc(fall3d_ads_ctx *ctx, double dt, double stime, int iiter);

/* Enqueue nsteps steps starting at iteration iiter0 with fixed dt/stime.
 * Step pairs are replayed from a captured hipGraph (re-captured when dt,
 * stime, the source or the tiling change) unless graphs are switched off
 * or profiling is on. */
int fall3d_ads_run(fall3d_ads_ctx *ctx, double dt, double stime, int iiter0, int nsteps);

/* Use (1, default) or bypass (0) graph replay in fall3d_ads_run. */
int fall3d_ads_set_graph(fall3d_ads_ctx *ctx, int on);

/* Reference flux accumulators W,E,S,N,D,U (FP64, device resident). */
int fall3d_ads_get_fluxes(fall3d_ads_ctx *ctx, double flux[6]);
int fall3d_ads_set_fluxes(fall3d_ads_ctx *ctx, const double flux[6]);

/* FP64 budget, FALL3D_ADS_BUDGET_* indices (synchronizes). */
int fall3d_ads_get_budget(fall3d_ads_ctx *ctx, double out[FALL3D_ADS_BUDGET_LEN]);

/* Restart (the device-side counterpart of nc_IO_out_rst / nc_IO_read_rst,
 * Sources/mod_nc_IO.F90:4524-5178, which store c with its halos, acum and
 * the budget attributes). The snapshot is a byte blob: a 64-byte header
 * (magic "F3DADSR1", version 1, int32 nx, ny, nz, nbins, state mode,
 * limiter, time marching, double time, int64 iiter, uint64 number of
 * accumulators), the FP64 accumulators (reference fluxes W,E,S,N,D,U,
 * ledger, injected mass, rounding residuals), acum (FP64, its reference
 * layout) and the device state words (FP32 c with ghosts; for the
 * compensated state also its low words). Restoring it into a context of the
 * same grid and state mode makes the following steps bitwise those of the
 * uninterrupted run. buf NULL: *bytes receives the size. time and iiter are
 * stored for the caller (restore returns them; either pointer may be NULL).
 * No netCDF: the reference's restart files need netCDF-Fortran. */
int fall3d_ads_snapshot(fall3d_ads_ctx *ctx, void *buf, size_t *bytes, double time, long long iiter);
int fall3d_ads_restore(fall3d_ads_ctx *ctx, const void *buf, size_t bytes, double *time, long long *iiter);

/* Health of the state, for checks after logged steps: counts[0] receives
 * the number of non-finite interior cells (all bins), counts[1] the number
 * of negative ones, min_value (may be NULL) the smallest finite value.
 * No counterpart in the reference (its runs stop on I/O of NaN values
 * only). Synchronizes. */
int fall3d_ads_check_state(fall3d_ads_ctx *ctx, long long counts[2], double *min_value);

/* Stream control: use an external hipStream_t (NULL = the library's own). */
int fall3d_ads_set_stream(fall3d_ads_ctx *ctx, void *stream);
int fall3d_ads_synchronize(fall3d_ads_ctx *ctx);

/* Tuning of the sweep of axis dir (0 x, 1 y, 2 z), 0 = automatic:
 * cells_per_thread is the x sweep's cells per lane (2, 3, 4, 6, 8, 12, 16
 * or 20; an x line of nx cells takes nx/C + 1 lanes, at most 256: one wave
 * holds one or more lines, or 2-4 waves of a workgroup hold one), and the
 * y/z sweeps' bins per thread (1 or 2; 2 not with the compensated state);
 * columns: x-columns per workgroup; bins_per_group: bins per workgroup. */
int fall3d_ads_set_tiling(fall3d_ads_ctx *ctx, int dir, int cells_per_thread, int columns, int bins_per_group);
/* The tile shape in use for axis dir: cells (x) or bins (y/z) per thread,
 * bins per workgroup. */
int fall3d_ads_get_tiling(fall3d_ads_ctx *ctx, int dir, int *cells_per_thread, int *bins_per_group);
/* Autotuner: times every candidate tile shape of each sweep (x: cells per
 * lane; y, z: bins per workgroup and bins per thread) on the context's
 * fields, keeps the fastest and restores state, deposit and accumulators.
 * tiling (3 ints: x cells per lane, y and z bins per workgroup; the rest via
 * fall3d_ads_get_tiling) and best_ms (3 doubles) may be NULL. */
int fall3d_ads_autotune(fall3d_ads_ctx *ctx, double dt, double stime, int reps, int tiling[3], double best_ms[3]);

/* Per-kernel timing with HIP events on the context stream; resets the totals.
 * on = 0: off; 1: every step (fall3d_ads_run then launches directly, no
 * graph); N > 1: fall3d_ads_run times the steps iiter0, iiter0 + N, ...
 * (launched directly) and replays its graph in between, so the events cost
 * one step in N. Single fall3d_ads_step calls are timed whenever on != 0.
 * kernel ids: 0 source, 1 sweep x, 2 sweep y, 3 sweep z, 4 reduce. */
int fall3d_ads_set_profiling(fall3d_ads_ctx *ctx, int on);
int fall3d_ads_get_kernel_time(fall3d_ads_ctx *ctx, int kernel, double *total_ms, long long *launches);

/* Algorithmic HBM bytes one launch of sweep `dir` moves (state + coefficient
 * arrays it must read once), for roofline accounting. */
int fall3d_ads_sweep_bytes(fall3d_ads_ctx *ctx, int dir, double *bytes);

/* Measurement aid, no counterpart in the reference: copies n floats from
 * src to dst (device pointers) with 16-byte buffer loads and stores, the
 * access width of the sweeps, on stream (NULL: the null stream), and
 * synchronizes. A kernel of known traffic (4n bytes read, 4n written) for
 * calibrating the HBM byte counters. */
int fall3d_ads_copy_probe(const void *src, void *dst, size_t n, void *stream);

/* Measurement aid, no counterpart in the reference: copies n doubles from
 * src to dst (device pointers) with 8-byte loads and stores, the access
 * width of the FP64 state's sweeps, on stream (NULL: the null stream), and
 * synchronizes. The known-traffic kernel (8n bytes read, 8n written) for
 * calibrating the HBM byte counters at that width. */
int fall3d_ads_copy_probe64(const void *src, void *dst, size_t n, void *stream);

/* ---- 3. Ensemble group --------------------------------------------------
 * FALL3D runs ensemble members (Sources/mod_Ensemble.F90, task_SetEns) as
 * independent task_Fall3d copies, one per MPI communicator
 * (Sources/mod_Parallel.F90:188-196), each with its own grid, perturbed
 * source and wind, and its own time step. An ensemble here holds one
 * context per member (sizes may differ; nbins, limiter, time marching and
 * state mode are common) and advances all of them with one batched launch
 * per sweep; a run of n steps is replayed from one captured hipGraph.
 * Members are filled and read through the resident entry points above on
 * the context returned by fall3d_ads_ensemble_member (borrowed: do not
 * release it, do not change its stream). No counterpart in the reference. */
typedef struct fall3d_ads_ensemble fall3d_ads_ensemble;

/* dims: 3*nmembers ints (nx, ny, nz of each member). */
int fall3d_ads_ensemble_create(fall3d_ads_ensemble **ens, int nmembers, const int *dims, int nbins, int limiter,
                               int time_marching, int real_kind, int state_mode);
int fall3d_ads_ensemble_release(fall3d_ads_ensemble *ens);
const char *fall3d_ads_ensemble_last_error(const fall3d_ads_ensemble *ens);
int fall3d_ads_ensemble_member(fall3d_ads_ensemble *ens, int m, fall3d_ads_ctx **ctx);
/* One F3D_time_step of every member; dt[m] and stime[m] per member (stime
 * may be NULL: 0 for all). Same order of work as fall3d_ads_step. */
int fall3d_ads_ensemble_step(fall3d_ads_ensemble *ens, const double *dt, const double *stime, int iiter);
/* nsteps steps from iteration iiter0 (graph replay; asynchronous). */
int fall3d_ads_ensemble_run(fall3d_ads_ensemble *ens, const double *dt, const double *stime, int iiter0,
                            int nsteps);
int fall3d_ads_ensemble_set_graph(fall3d_ads_ensemble *ens, int on);
int fall3d_ads_ensemble_synchronize(fall3d_ads_ensemble *ens);
/* Common tile shape: x-sweep cells per lane, y/z bins per workgroup (0 keeps
 * the current value). get_tiling returns the three values. */
int fall3d_ads_ensemble_set_tiling(fall3d_ads_ensemble *ens, int x_cells_per_lane, int y_bins_per_group,
                                   int z_bins_per_group);
int fall3d_ads_ensemble_get_tiling(fall3d_ads_ensemble *ens, int out[3]);
/* Autotuner: times every candidate tile shape of each sweep (reps launches
 * each) on the members as they are, keeps the fastest, restores the state
 * and accumulators. best_ms (3 doubles, may be NULL) receives the per-sweep
 * times of the chosen shapes. */
int fall3d_ads_ensemble_autotune(fall3d_ads_ensemble *ens, const double *dt, const double *stime, int reps,
                                 double *best_ms);

/* ---- 3a. Ensemble statistics (Sources/task_PosEns.F90:319-441) ----------
 * Statistics over the members of an ensemble at every grid point: x holds
 * nmem member fields of npoints values, member-major (x[m npoints + p]),
 * FP64, on the device (1 <= nmem <= 1024). Every output is a device array
 * and may be NULL (not computed):
 *   mean[p]                 sum_m x / nmem
 *   logmean[p]              10^(sum_m log10 x / nmem)
 *   stddev[p]               sqrt(sum_m (x - mean)^2 / (nmem - 1)) (the sum
 *                           the reference's loop at :356-360 is written for;
 *                           that loop keeps only the last member's term)
 *   median[p]               r(int(nmem/2)), r the member values in ascending
 *                           order (postp_rank_vector, Sources/mod_Postp.F90:185), 1-based
 *   per[i npoints + p]      r(min(max(int(percentiles[i]/100 nmem), 1), nmem))
 *   prob[i npoints + p]     100 #{m : x_m >= thresholds[i]} / nmem (percent)
 * percentiles (nper values, in percent) and thresholds (nth) are device
 * arrays. Runs on stream (NULL: the null stream) and synchronizes. */
int fall3d_ens_statistics(const double *x, int nmem, long npoints, const double *percentiles, int nper,
                          const double *thresholds, int nth, double *mean, double *logmean, double *stddev,
                          double *median, double *per, double *prob, void *stream);

/* ---- 4. Learned flux correction (coarse grid) ----------------------------
 * No counterpart in the reference. A per-face network (32 features -> 64 ->
 * 64 -> 1, GELU, BF16 weights and activations with FP32 accumulation on the
 * matrix cores, FP32 Adam) corrects the x and y face mass fluxes of a grid
 * coarsened 2x horizontally. Plain KT is the method and the fallback: a
 * context runs it unchanged until fall3d_ads_set_learned gives it a network.
 * Parameters are FP32 in the order W1[32][64], b1[64], W2[64][64], b2[64],
 * w3[64], b3 (fall3d_mlp_nparam() values). Feature rows are 32 bfloat16 bit
 * patterns (uint16). */
typedef struct fall3d_mlp fall3d_mlp;

/* capacity: training rows the network keeps on the device (0: none). */
int fall3d_mlp_create(fall3d_mlp **net, long capacity);
int fall3d_mlp_release(fall3d_mlp *net);
const char *fall3d_mlp_last_error(const fall3d_mlp *net);
int fall3d_mlp_nparam(void);
/* Set (also resets the Adam moments) / read the FP32 parameters. */
int fall3d_mlp_set_params(fall3d_mlp *net, const float *params);
int fall3d_mlp_get_params(fall3d_mlp *net, float *params);
/* y[r] = network(x[r]), host arrays. */
int fall3d_mlp_forward(fall3d_mlp *net, const uint16_t *x, long rows, float *y);
/* Gradient of the mean squared error 1/rows sum (y - t)^2 (host arrays);
 * sse (may be NULL) receives sum (y - t)^2. */
int fall3d_mlp_gradient(fall3d_mlp *net, const uint16_t *x, const float *t, long rows, float *grad, double *sse);
/* Training rows: append host rows, count, read back, clear. */
int fall3d_mlp_data_append(fall3d_mlp *net, const uint16_t *x, const float *t, long rows);
int fall3d_mlp_data_count(fall3d_mlp *net, long *rows);
int fall3d_mlp_data_get(fall3d_mlp *net, long first, long rows, uint16_t *x, float *t);
int fall3d_mlp_data_clear(fall3d_mlp *net);
/* Rows collected by fall3d_ads_learned_collect: face of each row (4 int32:
 * axis | bin << 1, face index f, cross index, level k) and sgn S dxp. */
int fall3d_mlp_data_refs(fall3d_mlp *net, long first, long rows, int32_t *ref, float *den);
/* steps Adam updates (lr; beta 0.9, 0.999; eps 1e-8) on uniform random
 * batches of the stored rows; loss (steps floats, may be NULL) receives the
 * mean squared error of each batch. */
int fall3d_mlp_train(fall3d_mlp *net, int steps, int batch, float lr, unsigned long long seed, float *loss);

/* x and y face mass flux per unit area of every later sweep, accumulated
 * (minmod, RK4, FP32 state); the call zeroes the accumulators. get copies
 * axis 0 (u layout x nbins) or 1 (v layout x nbins) to the host. */
int fall3d_ads_set_flux_output(fall3d_ads_ctx *ctx, int on);
int fall3d_ads_get_face_flux(fall3d_ads_ctx *ctx, int axis, float *out);
/* Use net (borrowed; NULL: plain KT) in every later step. Faces whose
 * stencil maximum is <= smin get no correction (smin < 0: <= |smin| times the
 * field's largest value, found on the device each step); gain scales the
 * applied correction (1: as predicted); flux_limit > 0 bounds each face's
 * correction by flux_limit times the plain step's flux through that face. */
int fall3d_ads_set_learned(fall3d_ads_ctx *ctx, fall3d_mlp *net, double smin, double gain, double flux_limit);
/* Plain KT is the fallback of the learned step: a step in which the network
 * gives a non-finite output for any active face applies no correction at
 * all (the state is then bitwise the plain KT step). steps receives the
 * number of such steps since fall3d_ads_set_learned. */
int fall3d_ads_learned_fallbacks(fall3d_ads_ctx *ctx, long long *steps);
/* Interior of coarse := 2x2 horizontal means of fine (fine = 2x coarse). */
int fall3d_ads_coarse_grain(fall3d_ads_ctx *fine, fall3d_ads_ctx *coarse);
/* One training sample of the whole coarse grid: coarse := coarse-grained
 * fine; features of the active coarse faces (stencil maximum > smin); one
 * plain coarse step (dt) and two fine steps (dt/2, advancing fine); targets
 * from the face fluxes. Appends the rows to net; added receives their count. */
int fall3d_ads_learned_collect(fall3d_ads_ctx *fine, fall3d_ads_ctx *coarse, fall3d_mlp *net, double dt,
                               double stime, int iiter, double smin, long *added);

#ifdef __cplusplus
}
#endif

#endif /* FALL3D_ADS_H */
