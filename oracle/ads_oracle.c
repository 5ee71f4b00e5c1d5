/*
 * ads_oracle.c -- FP64 CPU restatement of FALL3D's advection-diffusion-
 * sedimentation (ADS) step, used ONLY as the test/benchmark checker.
 *
 * TEST INFRASTRUCTURE. Nothing in the product path (the HIP library under
 * fall3d_amd/) links, loads or calls this file. It is imported only by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
 *
 * What it restates (reference FALL3D 8.2.1, paths relative to the tree root):
 *   Sources/mod_ADS.F90:233-412   ADS_solve_along_x   -> ora_solve_x
 *   Sources/mod_ADS.F90:422-602   ADS_solve_along_y   -> ora_solve_y
 *   Sources/mod_ADS.F90:612-811   ADS_solve_along_z   -> ora_solve_z
 *   Sources/mod_ADS.F90:1057-1120 KT_RHS              -> kt_rhs
 *   Sources/mod_ADS.F90:1124-1223 r_minmod/r_sbee/r_ospre -> limiter
 *   Sources/mod_ADS.F90:1227-1326 freeflow            -> freeflow_*
 *   Sources/mod_F3D.F90:1456-1491 source add + sweep order -> ora_step
 *   Sources/mod_Grid.F90:1026-1141 grid_get_time_step -> ora_time_step
 *   Sources/mod_Grid.F90:1340-1380 grid_get_mass_volume -> ora_mass_volume
 *
 * Array layouts are the reference's (Fortran column-major, explicit bounds
 * from Sources/mod_Domain.F90:263-330 for a serial run: ips=1, ipe=nx,
 * ips_2h=-1, ipe_2h=nx+2, ibs_1h=0, ibe_1h=nx+2, same for j and k):
 *   c    (-1:nx+2, -1:ny+2, -1:nz+2, nb)     x fastest
 *   u    ( 0:nx+2,  1:ny,    1:nz)           x fastest (face i = west face of cell i)
 *   v    ( 0:ny+2,  1:nx,    1:nz)           y fastest
 *   w    ( 0:nz+2,  1:nx,    1:ny)           z fastest
 *   k1   (-1:nx+2,  1:ny,    1:nz)           k2 (-1:ny+2,1:nx,1:nz)  k3 (-1:nz+2,1:nx,1:ny)
 *   vs   ( 1:nx,    1:ny,    0:nz+2, nb)
 *   acum (-1:nx+2, -1:ny+2, nb)
 *   s    ( 1:nx,    1:ny,    1:nz,   nb)
 *   grid dX1_p(-1:nx+2) dX1_b(0:nx+2) and likewise for 2 and 3.
 *
 * The arithmetic follows the reference statement by statement (same operand
 * order, no FMA contraction: build with -ffp-contract=off) so that it tracks
 * the reference to rounding level. Line loops run under OpenMP; boundary
 * sums are reduced serially in the reference's loop order, so results do not
 * depend on the thread count.
 *
 * Besides the reference-compatible boundary fluxes (last RK stage only,
 * advective + diffusive with the reference's "+P" sign, mod_ADS.F90:1116-1117,
 * 394-396) each solver can also fill a conservation ledger: the mass that
 * actually left through each end face, i.e. dt*area*sum_s w_s (F_s - P_s)
 * with the RK4 weights 1/6,1/3,1/3,1/6. This is not a reference quantity;
 * it is what the telescoping flux form conserves exactly.
 */
#include <math.h>
#include <stdlib.h>
#include <stddef.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORA_EPSILON 1e-12 /* Sources/mod_KindType.F90:42 */

enum { LIM_MINMOD = 1, LIM_SUPERBEE = 2, LIM_OSPRE = 3 };
enum { TM_EULER = 1, TM_RK = 2 };

/* cfg[] layout shared with the Python wrapper */
enum {
  CFG_NX = 0, CFG_NY, CFG_NZ, CFG_NB,
  CFG_BW, CFG_BE, CFG_BS, CFG_BN, CFG_BD, CFG_BU,
  CFG_LIMITER, CFG_TM,
  CFG_BCMODE, /* 0 freeflow (the reference's choice, mod_ADS.F90:116), 1 dirichlet (:1330-1399),
                 2 periodic (:1403-1448, in the reference a zero-gradient copy) */
  CFG_LEN
};

/* ---------------------------------------------------------------------- */
/* 1-D kernels. A line is passed as base pointers shifted so that the      */
/* reference's Fortran index can be used directly: c[i] for i in -1..n+2,  */
/* dxp[i] for i in -1..n+2, dxb[i] and u[i] for i in 0..n+2, k[i] in -1..n+2 */
/* ---------------------------------------------------------------------- */

/* Sources/mod_ADS.F90:1124-1223. r is indexed -1..n+2 (r[-1] and r[n+2]
 * are copies of the neighbours, as in the reference). */
static void limiter(int lim, int n, const double *c, double *r) {
  for (int i = 0; i <= n + 1; ++i) {
    double dp = c[i + 1] - c[i];
    if (fabs(dp) < ORA_EPSILON) {
      if (lim == LIM_MINMOD) {
        r[i] = 1.0;
      } else {
        double dm = c[i] - c[i - 1];
        if (fabs(dm) < ORA_EPSILON)
          r[i] = 0.0;
        else
          r[i] = (lim == LIM_SUPERBEE) ? 2.0 : 1.5;
      }
    } else {
      double rr = (c[i] - c[i - 1]) / (c[i + 1] - c[i]);
      if (lim == LIM_MINMOD) {
        double t = fmin(1.0, rr);
        r[i] = fmax(0.0, t);
      } else if (lim == LIM_SUPERBEE) {
        double a = fmin(2.0 * rr, 1.0);
        double b = fmin(rr, 2.0);
        double m = fmax(0.0, a);
        r[i] = fmax(m, b);
      } else {
        r[i] = 1.5 * (rr * rr + rr) / (rr * rr + rr + 1.0);
      }
    }
  }
  r[-1] = r[0];
  r[n + 2] = r[n + 1];
}

/* Sources/mod_ADS.F90:1057-1120. Work arrays: r (n+4), cr/cl (n+2). The
 * advective (F) and diffusive (P) end-face fluxes are also returned so the
 * caller can form both the reference flux (F+P) and the ledger (F-P). */
static void kt_rhs(int lim, int n, const double *c, const double *dxp,
                   const double *dxb, const double *u, const double *k,
                   double *rhs /* 1..n */, double *Fends /*[2]*/,
                   double *Pends /*[2]*/, double *rbuf, double *crbuf,
                   double *clbuf) {
  double *r = rbuf + 1;   /* r[-1..n+2]  */
  double *cr = crbuf + 0; /* cr[0..n+1]  */
  double *cl = clbuf + 0;
  limiter(lim, n, c, r);
  for (int i = 0; i <= n + 1; ++i) {
    cr[i] = c[i] - 0.5 * r[i] * (c[i + 1] - c[i]);
    cl[i] = c[i - 1] + 0.5 * r[i - 1] * (c[i] - c[i - 1]);
  }
  for (int i = 1; i <= n; ++i) {
    double F_ml = cl[i] * u[i];
    double F_mr = cr[i] * u[i];
    double F_pl = cl[i + 1] * u[i + 1];
    double F_pr = cr[i + 1] * u[i + 1];
    double F_m = 0.5 * (F_ml + F_mr) - 0.5 * fabs(u[i]) * (cr[i] - cl[i]);
    double F_p = 0.5 * (F_pl + F_pr) - 0.5 * fabs(u[i + 1]) * (cr[i + 1] - cl[i + 1]);
    double P_m = 0.5 * (k[i] + k[i - 1]) * (c[i] - c[i - 1]) / dxp[i - 1];
    double P_p = 0.5 * (k[i + 1] + k[i]) * (c[i + 1] - c[i]) / dxp[i];
    rhs[i] = -1.0 * (F_p - F_m) / dxb[i] + (P_p - P_m) / dxb[i];
    if (i == 1) { Fends[0] = F_m; Pends[0] = P_m; }
    if (i == n) { Fends[1] = F_p; Pends[1] = P_p; }
  }
}

/* Per-thread line workspace. */
typedef struct {
  double *line, *c0, *nc, *rhs, *r, *cr, *cl, *uu, *kk;
} lws;

static void lws_alloc(lws *w, int nmax) {
  int m = nmax + 8;
  w->line = calloc(m, sizeof(double));
  w->c0 = calloc(m, sizeof(double));
  w->nc = calloc(m, sizeof(double));
  w->rhs = calloc(m, sizeof(double));
  w->r = calloc(m, sizeof(double));
  w->cr = calloc(m, sizeof(double));
  w->cl = calloc(m, sizeof(double));
  w->uu = calloc(m, sizeof(double));
  w->kk = calloc(m, sizeof(double));
}
static void lws_free(lws *w) {
  free(w->line); free(w->c0); free(w->nc); free(w->rhs); free(w->r);
  free(w->cr); free(w->cl); free(w->uu); free(w->kk);
}

/* Integrate one line in place (line[-1..n+2], ghosts already set by the
 * boundary condition). Stage structure of mod_ADS.F90:288-399: the first
 * stage copies the ghosts into c0/new_c (:330,:333); later stages update the
 * interior only, so ghosts stay frozen for the whole sweep.
 * flux_ref[2] receives the reference end-face fluxes of the last stage,
 * ledger[2] the RK-weighted end-face net fluxes (F-P). */
static void integrate_line(int lim, int tm, int n, double dt, double *line,
                           const double *dxp, const double *dxb,
                           const double *u, const double *k, lws *w,
                           double flux_ref[2], double ledger[2]) {
  double Fe[2], Pe[2];
  double *rhs = w->rhs;
  double *c0 = w->c0 + 1;
  double *nc = w->nc + 1;
  if (tm == TM_EULER) {
    kt_rhs(lim, n, line, dxp, dxb, u, k, rhs, Fe, Pe, w->r, w->cr, w->cl);
    for (int i = 1; i <= n; ++i) line[i] = line[i] + dt * rhs[i];
    flux_ref[0] = Fe[0] + Pe[0];
    flux_ref[1] = Fe[1] + Pe[1];
    ledger[0] = Fe[0] - Pe[0];
    ledger[1] = Fe[1] - Pe[1];
    return;
  }
  /* stage 1 */
  kt_rhs(lim, n, line, dxp, dxb, u, k, rhs, Fe, Pe, w->r, w->cr, w->cl);
  double g0 = (Fe[0] - Pe[0]) / 6.0, g1 = (Fe[1] - Pe[1]) / 6.0;
  for (int i = -1; i <= n + 2; ++i) { c0[i] = line[i]; nc[i] = line[i]; }
  for (int i = 1; i <= n; ++i) {
    c0[i] = c0[i] + 0.5 * dt * rhs[i];
    nc[i] = nc[i] + dt * rhs[i] / 6.0;
  }
  /* stage 2 */
  kt_rhs(lim, n, c0, dxp, dxb, u, k, rhs, Fe, Pe, w->r, w->cr, w->cl);
  g0 += (Fe[0] - Pe[0]) / 3.0; g1 += (Fe[1] - Pe[1]) / 3.0;
  for (int i = 1; i <= n; ++i) {
    c0[i] = line[i] + 0.5 * dt * rhs[i];
    nc[i] = nc[i] + dt * rhs[i] / 3.0;
  }
  /* stage 3 */
  kt_rhs(lim, n, c0, dxp, dxb, u, k, rhs, Fe, Pe, w->r, w->cr, w->cl);
  g0 += (Fe[0] - Pe[0]) / 3.0; g1 += (Fe[1] - Pe[1]) / 3.0;
  for (int i = 1; i <= n; ++i) {
    c0[i] = line[i] + dt * rhs[i];
    nc[i] = nc[i] + dt * rhs[i] / 3.0;
  }
  /* stage 4 */
  kt_rhs(lim, n, c0, dxp, dxb, u, k, rhs, Fe, Pe, w->r, w->cr, w->cl);
  g0 += (Fe[0] - Pe[0]) / 6.0; g1 += (Fe[1] - Pe[1]) / 6.0;
  for (int i = 1; i <= n; ++i) line[i] = nc[i] + dt * rhs[i] / 6.0;
  flux_ref[0] = Fe[0] + Pe[0];
  flux_ref[1] = Fe[1] + Pe[1];
  ledger[0] = g0;
  ledger[1] = g1;
}

/* Boundary ghosts for one gathered line. Free flow (mode 0,
 * Sources/mod_ADS.F90:1239-1324): inflow (u_first > 0 / u_last < 0) gives
 * 2*bvalue - c with bvalue = 0 (:117), otherwise a zero-gradient copy.
 * Dirichlet (mode 1, :1330-1399): 2*bvalue - c on every physical side.
 * Periodic (mode 2, :1403-1448): a zero-gradient copy on every side. The
 * outer ghost copies the inner one. */
static void freeflow_line(int n, double *line, double ubeg, double uend,
                          int bbeg, int bend, int mode) {
  if (bbeg) {
    const int in = mode == 0 ? (ubeg > 0.0) : (mode == 1);
    if (in) line[0] = 2.0 * 0.0 - line[1];
    else line[0] = line[1];
    line[-1] = line[0];
  }
  if (bend) {
    const int in = mode == 0 ? (uend < 0.0) : (mode == 1);
    if (in) line[n + 1] = 2.0 * 0.0 - line[n];
    else line[n + 1] = line[n];
    line[n + 2] = line[n + 1];
  }
}

#define CIDX(i, j, k, b) \
  ((size_t)((i) + 1) + (size_t)(nx + 4) * ((size_t)((j) + 1) + (size_t)(ny + 4) * ((size_t)((k) + 1) + (size_t)(nz + 4) * (size_t)(b))))

static int max3(int a, int b, int c) { int m = a > b ? a : b; return m > c ? m : c; }

/* ---------------------------------------------------------------------- */
/* Sweeps                                                                   */
/* ---------------------------------------------------------------------- */

/* dir: 0 = x, 1 = y, 2 = z. For z, `vel` is w1, vel2 = w2, stime and vs are
 * used to form w_eff = (1-s) w1 + s w2 - vs (mod_ADS.F90:667-672). */
static void sweep(int dir, const int *cfg, const double *const *grid,
                  double dt, double *c, const double *kdiff, const double *vel,
                  const double *vel2, double stime, const double *vs,
                  double *acum, double *fbeg, double *fend, double *ledger) {
  const int nx = cfg[CFG_NX], ny = cfg[CFG_NY], nz = cfg[CFG_NZ], nb = cfg[CFG_NB];
  const int lim = cfg[CFG_LIMITER], tm = cfg[CFG_TM];
  const int bbeg = cfg[CFG_BW + 2 * dir], bend = cfg[CFG_BE + 2 * dir];
  const int n = dir == 0 ? nx : (dir == 1 ? ny : nz);
  const int n1 = dir == 0 ? ny : nx;          /* first transverse index range */
  const int n2 = dir == 2 ? ny : nz;          /* second transverse index range */
  const double *dxp = grid[2 * dir] + 1;      /* dX_p(-1:n+2): position p at index p+1 */
  const double *dxb = grid[2 * dir + 1];      /* dX_b(0:n+2): position p at index p */
  /* transverse cell sizes for the face area */
  const double *dX1b = grid[1], *dX2b = grid[3], *dX3b = grid[5];
  const size_t nlines = (size_t)n1 * n2;
  double *fl = malloc(sizeof(double) * 4 * nlines * nb); /* ref flux beg/end, ledger beg/end */

#pragma omp parallel
  {
    lws w;
    lws_alloc(&w, max3(nx, ny, nz));
    double *line = w.line + 2; /* line[-1..n+2] */
    double *uu = w.uu + 1;     /* uu[0..n+2] */
    double *kk = w.kk + 2;     /* kk[-1..n+2] */
#pragma omp for collapse(2) schedule(static)
    for (int b = 0; b < nb; ++b) {
      for (long L = 0; L < (long)nlines; ++L) {
        int t1 = (int)(L % n1) + 1, t2 = (int)(L / n1) + 1;
        /* reference line loops: x: k outer, j inner; y: k outer, i inner;
         * z: j outer, i inner. (t1,t2) = (j,k), (i,k), (i,j) respectively. */
        size_t stride;
        size_t base;
        double area;
        if (dir == 0) {
          base = CIDX(0, t1, t2, b); stride = 1;
          for (int i = 0; i <= n + 2; ++i) uu[i] = vel[(size_t)i + (size_t)(nx + 3) * ((t1 - 1) + (size_t)ny * (t2 - 1))];
          for (int i = -1; i <= n + 2; ++i) kk[i] = kdiff[(size_t)(i + 1) + (size_t)(nx + 4) * ((t1 - 1) + (size_t)ny * (t2 - 1))];
          area = dX3b[t2] * dX2b[t1];
        } else if (dir == 1) {
          base = CIDX(t1, 0, t2, b); stride = (size_t)(nx + 4);
          for (int j = 0; j <= n + 2; ++j) uu[j] = vel[(size_t)j + (size_t)(ny + 3) * ((t1 - 1) + (size_t)nx * (t2 - 1))];
          for (int j = -1; j <= n + 2; ++j) kk[j] = kdiff[(size_t)(j + 1) + (size_t)(ny + 4) * ((t1 - 1) + (size_t)nx * (t2 - 1))];
          area = dX3b[t2] * dX1b[t1];
        } else {
          base = CIDX(t1, t2, 0, b); stride = (size_t)(nx + 4) * (ny + 4);
          for (int k = 0; k <= n + 2; ++k) {
            size_t wi = (size_t)k + (size_t)(nz + 3) * ((t1 - 1) + (size_t)nx * (t2 - 1));
            double wv = (1.0 - stime) * vel[wi] + stime * vel2[wi];
            size_t vi = (size_t)(t1 - 1) + (size_t)nx * ((t2 - 1) + (size_t)ny * ((size_t)k + (size_t)(nz + 3) * b));
            uu[k] = wv - vs[vi];
          }
          for (int k = -1; k <= n + 2; ++k) kk[k] = kdiff[(size_t)(k + 1) + (size_t)(nz + 4) * ((t1 - 1) + (size_t)nx * (t2 - 1))];
          area = dX2b[t2] * dX1b[t1];
        }
        /* gather (ghosts included), apply BC, integrate, scatter */
        for (int i = -1; i <= n + 2; ++i) line[i] = c[(ptrdiff_t)base + (ptrdiff_t)i * (ptrdiff_t)stride];
        freeflow_line(n, line, uu[1], uu[n + 1], bbeg, bend, cfg[CFG_BCMODE]);
        double fr[2], lg[2];
        integrate_line(lim, tm, n, dt, line, dxp, dxb, uu, kk, &w, fr, lg);
        for (int i = -1; i <= n + 2; ++i) c[(ptrdiff_t)base + (ptrdiff_t)i * (ptrdiff_t)stride] = line[i];
        size_t o = 4 * ((size_t)b * nlines + (size_t)L);
        fl[o + 2] = lg[0] * area * dt;
        fl[o + 3] = lg[1] * area * dt;
        if (dir == 2 && bbeg && acum) {
          size_t ai = (size_t)(t1 + 1) + (size_t)(nx + 4) * ((size_t)(t2 + 1) + (size_t)(ny + 4) * b);
          acum[ai] = acum[ai] + fabs(fr[0]) * dt;
        }
        fl[o + 0] = fr[0] * area * dt;
        fl[o + 1] = fr[1] * area * dt;
      }
    }
    lws_free(&w);
  }
  /* serial reduction in the reference's order (bin, outer line, inner line) */
  for (int b = 0; b < nb; ++b) {
    for (size_t L = 0; L < nlines; ++L) {
      size_t o = 4 * ((size_t)b * nlines + L);
      if (bbeg) *fbeg = *fbeg + fl[o + 0];
      if (bend) *fend = *fend + fl[o + 1];
      if (ledger) {
        if (bbeg) ledger[0] -= fl[o + 2]; /* mass leaving through the first face */
        if (bend) ledger[1] += fl[o + 3]; /* mass leaving through the last face  */
      }
    }
  }
  free(fl);
}

/* c: base of c(-1,-1,-1,1). The base pointer of the gathered line uses
 * index i directly (CIDX(0,...) is the i=0 element), so line[i] maps to
 * c[base + i*stride] for i in -1..n+2. */

void ora_solve_x(const int *cfg, const double *const *grid, double dt,
                 double *c, const double *k1, const double *u, double *Wf,
                 double *Ef, double *ledger) {
  sweep(0, cfg, grid, dt, c, k1, u, NULL, 0.0, NULL, NULL, Wf, Ef, ledger);
}

void ora_solve_y(const int *cfg, const double *const *grid, double dt,
                 double *c, const double *k2, const double *v, double *Sf,
                 double *Nf, double *ledger) {
  sweep(1, cfg, grid, dt, c, k2, v, NULL, 0.0, NULL, NULL, Sf, Nf, ledger);
}

void ora_solve_z(const int *cfg, const double *const *grid, double dt,
                 double *c, double *acum, const double *k3, const double *w1,
                 const double *w2, double stime, const double *vs, double *Df,
                 double *Uf, double *ledger) {
  sweep(2, cfg, grid, dt, c, k3, w1, w2, stime, vs, acum, Df, Uf, ledger);
}

/* Sources/mod_F3D.F90:1456-1473: c += dt*s/vol with
 * vol = dX1_p(i) dX2_p(j) dX3_p(k) Hm1_p(j) Hm2_p(j) Hm3_p(i,j). The map
 * factors are optional: hm1(1:ny), hm2(-1:ny+2), hm3(1:nx,1:ny), NULL = 1
 * (flat Cartesian grid, Sources/mod_Grid.F90:403,463,603). Returns the
 * injected mass sum(dt*s/(Hm1 Hm2 Hm3)) in the dX_p measure. */
double ora_add_source_metric(const int *cfg, const double *const *grid, double dt,
                             double *c, const double *s, const double *hm1,
                             const double *hm2, const double *hm3) {
  const int nx = cfg[CFG_NX], ny = cfg[CFG_NY], nz = cfg[CFG_NZ], nb = cfg[CFG_NB];
  const double *dx1p = grid[0] + 1, *dx2p = grid[2] + 1, *dx3p = grid[4] + 1;
  double injected = 0.0;
  for (int k = 1; k <= nz; ++k) {
    double dZ = dx3p[k];
    for (int j = 1; j <= ny; ++j) {
      double dY = dx2p[j];
      double Hm1 = hm1 ? hm1[j - 1] : 1.0;
      double Hm2 = hm2 ? hm2[j + 1] : 1.0;
      for (int i = 1; i <= nx; ++i) {
        double dX = dx1p[i];
        double Hm3 = hm3 ? hm3[(size_t)(i - 1) + (size_t)nx * (size_t)(j - 1)] : 1.0;
        double vol = dX * dY * dZ * (Hm1 * Hm2 * Hm3);
        for (int b = 0; b < nb; ++b) {
          size_t si = (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * ((size_t)(k - 1) + (size_t)nz * b));
          c[CIDX(i, j, k, b)] = c[CIDX(i, j, k, b)] + dt * s[si] / vol;
          injected += dt * s[si] / (Hm1 * Hm2 * Hm3);
        }
      }
    }
  }
  return injected;
}

double ora_add_source(const int *cfg, const double *const *grid, double dt,
                      double *c, const double *s) {
  return ora_add_source_metric(cfg, grid, dt, c, s, NULL, NULL, NULL);
}

/* Sources/mod_Grid.F90:1026-1141, criterion 1 (per direction) or 2 (all
 * directions at once, default). Returns CFL/max(dtinv) (0 if no motion). */
double ora_time_step(const int *cfg, const double *const *grid,
                     const double *u, const double *v, const double *w,
                     const double *k1, const double *k2, const double *k3,
                     const double *vs, int criterion, double cfl) {
  const int nx = cfg[CFG_NX], ny = cfg[CFG_NY], nz = cfg[CFG_NZ], nb = cfg[CFG_NB];
  const double *dx1p = grid[0] + 1, *dx2p = grid[2] + 1, *dx3p = grid[4] + 1;
  double dt1 = 0.0, dt2 = 0.0, dt3 = 0.0;
#define U_(i, j, k) u[(size_t)(i) + (size_t)(nx + 3) * ((size_t)((j) - 1) + (size_t)ny * ((k) - 1))]
#define V_(j, i, k) v[(size_t)(j) + (size_t)(ny + 3) * ((size_t)((i) - 1) + (size_t)nx * ((k) - 1))]
#define W_(k, i, j) w[(size_t)(k) + (size_t)(nz + 3) * ((size_t)((i) - 1) + (size_t)nx * ((j) - 1))]
#define K1_(i, j, k) k1[(size_t)((i) + 1) + (size_t)(nx + 4) * ((size_t)((j) - 1) + (size_t)ny * ((k) - 1))]
#define K2_(j, i, k) k2[(size_t)((j) + 1) + (size_t)(ny + 4) * ((size_t)((i) - 1) + (size_t)nx * ((k) - 1))]
#define K3_(k, i, j) k3[(size_t)((k) + 1) + (size_t)(nz + 4) * ((size_t)((i) - 1) + (size_t)nx * ((j) - 1))]
#define VS_(i, j, k, b) vs[(size_t)((i) - 1) + (size_t)nx * ((size_t)((j) - 1) + (size_t)ny * ((size_t)(k) + (size_t)(nz + 3) * (b)))]
  for (int k = 1; k <= nz; ++k) {
    double dz = dx3p[k];
    for (int j = 1; j <= ny; ++j) {
      double dy = dx2p[j];
      for (int i = 1; i <= nx; ++i) {
        double dx = dx1p[i];
        double wmax = 0.0;
        for (int b = 0; b < nb; ++b) {
          double a = fabs(W_(k, i, j) - VS_(i, j, k, b));
          if (b == 0 || a > wmax) wmax = a;
        }
        if (criterion == 1) {
          double t1 = 2.0 * K1_(i, j, k) / (dx * dx) + fabs(U_(i, j, k)) / dx;
          double t2 = 2.0 * K2_(j, i, k) / (dy * dy) + fabs(V_(j, i, k)) / dy;
          double t3 = 2.0 * K3_(k, i, j) / (dz * dz) + wmax / dz;
          dt1 = fmax(dt1, t1); dt2 = fmax(dt2, t2); dt3 = fmax(dt3, t3);
        } else {
          double dtinv = 2.0 * K1_(i, j, k) / (dx * dx) + fabs(U_(i, j, k)) / dx +
                         2.0 * K2_(j, i, k) / (dy * dy) + fabs(V_(j, i, k)) / dy +
                         2.0 * K3_(k, i, j) / (dz * dz) + wmax / dz;
          dt1 = fmax(dt1, dtinv);
        }
      }
    }
  }
  double my_dt = fmax(fmax(dt1, dt2), dt3);
  if (my_dt > 0.0) my_dt = 1.0 / my_dt;
  return cfl * my_dt;
#undef U_
#undef V_
#undef W_
#undef K1_
#undef K2_
#undef K3_
#undef VS_
}

/* Sources/mod_Grid.F90:1364-1375 (mass-point spacings, bins innermost). */
double ora_mass_volume(const int *cfg, const double *const *grid, const double *c) {
  const int nx = cfg[CFG_NX], ny = cfg[CFG_NY], nz = cfg[CFG_NZ], nb = cfg[CFG_NB];
  const double *dx1p = grid[0] + 1, *dx2p = grid[2] + 1, *dx3p = grid[4] + 1;
  double m = 0.0;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i)
        for (int b = 0; b < nb; ++b)
          m = m + c[CIDX(i, j, k, b)] * dx1p[i] * dx2p[j] * dx3p[k];
  return m;
}

/* Mass in the conservative measure c*dX1_b*dX2_b*dX3_b (the measure the
 * telescoping flux form conserves; equal to the above on uniform grids). */
double ora_mass_cells(const int *cfg, const double *const *grid, const double *c) {
  const int nx = cfg[CFG_NX], ny = cfg[CFG_NY], nz = cfg[CFG_NZ], nb = cfg[CFG_NB];
  const double *dx1b = grid[1], *dx2b = grid[3], *dx3b = grid[5];
  double m = 0.0;
  for (int b = 0; b < nb; ++b)
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i)
          m += c[CIDX(i, j, k, b)] * dx1b[i] * dx2b[j] * dx3b[k];
  return m;
}

/* One F3D_time_step (Sources/mod_F3D.F90:1437-1491) for steady winds
 * (u = u1 = u2 so the velocity interpolation is the identity): source,
 * x/y order by iiter parity, then z. flux6: W,E,S,N,D,U reference
 * accumulators (in/out); ledger6 likewise for the conservation ledger.
 * Returns the injected mass. */
double ora_step(const int *cfg, const double *const *grid, double dt, int iiter,
                double *c, double *acum, const double *u, const double *v,
                const double *w1, const double *w2, double stime,
                const double *k1, const double *k2, const double *k3,
                const double *vs, const double *s, double *flux6,
                double *ledger6) {
  double inj = 0.0;
  if (s) inj = ora_add_source(cfg, grid, dt, c, s);
  if (iiter % 2 == 0) {
    ora_solve_x(cfg, grid, dt, c, k1, u, &flux6[0], &flux6[1], ledger6 ? &ledger6[0] : NULL);
    ora_solve_y(cfg, grid, dt, c, k2, v, &flux6[2], &flux6[3], ledger6 ? &ledger6[2] : NULL);
  } else {
    ora_solve_y(cfg, grid, dt, c, k2, v, &flux6[2], &flux6[3], ledger6 ? &ledger6[2] : NULL);
    ora_solve_x(cfg, grid, dt, c, k1, u, &flux6[0], &flux6[1], ledger6 ? &ledger6[0] : NULL);
  }
  ora_solve_z(cfg, grid, dt, c, acum, k3, w1, w2, stime, vs, &flux6[4], &flux6[5], ledger6 ? &ledger6[4] : NULL);
  return inj;
}


/* Sources/mod_ADS.F90:820-1053 ADS_solve_1D: one line, per-side boundary
 * code bcode in {0 none, -1 periodic (applied only for inflow at that side,
 * as in the reference), 1 Dirichlet, 2 free flow}; RK4 combined as
 * c + dt (k1 + 2 k2 + 2 k3 + k4)/6; flux = end-face fluxes of the last stage.
 * Arrays are indexed from the line's first halo element (c, dxp, k: n+4;
 * dxb, u: n+3). The reference's re-allocation of already allocated module
 * work arrays (:871-873) is not reproduced. */
void ora_solve_1d(int n, double *cl, const double *dxpl, const double *dxbl,
                  const double *ul, const double *kl, double dt,
                  const double *bvalue, double *flux, const int *bcode, int lim) {
  double *c = cl + 1;            /* c[-1..n+2] */
  const double *dxp = dxpl + 1;  /* dxp[-1..n+2] */
  const double *dxb = dxbl;      /* dxb[0..n+2] */
  const double *u = ul;          /* u[0..n+2]  */
  const double *k = kl + 1;
  lws w;
  lws_alloc(&w, n);
  double *c0 = w.c0 + 1;
  double *kk[4];
  for (int s = 0; s < 4; ++s) kk[s] = calloc((size_t)n + 2, sizeof(double));
  if (bcode[0] == -1) {
    if (u[1] > 0.0) { c[0] = c[n]; c[-1] = c[n - 1]; c[n + 1] = c[1]; c[n + 2] = c[2]; }
  } else if (bcode[0] == 1) {
    c[0] = 2.0 * bvalue[0] - c[1]; c[-1] = c[0];
  } else if (bcode[0] == 2) {
    if (u[1] > 0.0) { c[0] = 2.0 * bvalue[0] - c[1]; c[-1] = c[0]; }
    else { c[0] = c[1]; c[-1] = c[0]; }
  }
  if (bcode[1] == -1) {
    if (u[n + 1] < 0.0) { c[0] = c[n]; c[-1] = c[n - 1]; c[n + 1] = c[1]; c[n + 2] = c[2]; }
  } else if (bcode[1] == 1) {
    c[n + 1] = 2.0 * bvalue[1] - c[n]; c[n + 2] = c[n + 1];
  } else if (bcode[1] == 2) {
    if (u[n + 1] < 0.0) { c[n + 1] = 2.0 * bvalue[1] - c[n]; c[n + 2] = c[n + 1]; }
    else { c[n + 1] = c[n]; c[n + 2] = c[n + 1]; }
  }
  double Fe[2], Pe[2];
  for (int i = -1; i <= n + 2; ++i) c0[i] = c[i];
  kt_rhs(lim, n, c0, dxp, dxb, u, k, kk[0], Fe, Pe, w.r, w.cr, w.cl);
  for (int i = 1; i <= n; ++i) c0[i] = c[i] + 0.5 * dt * kk[0][i];
  kt_rhs(lim, n, c0, dxp, dxb, u, k, kk[1], Fe, Pe, w.r, w.cr, w.cl);
  for (int i = 1; i <= n; ++i) c0[i] = c[i] + 0.5 * dt * kk[1][i];
  kt_rhs(lim, n, c0, dxp, dxb, u, k, kk[2], Fe, Pe, w.r, w.cr, w.cl);
  for (int i = 1; i <= n; ++i) c0[i] = c[i] + dt * kk[2][i];
  kt_rhs(lim, n, c0, dxp, dxb, u, k, kk[3], Fe, Pe, w.r, w.cr, w.cl);
  for (int i = 1; i <= n; ++i)
    c[i] = c[i] + dt * (kk[0][i] + 2.0 * kk[1][i] + 2.0 * kk[2][i] + kk[3][i]) / 6.0;
  flux[0] = Fe[0] + Pe[0];
  flux[1] = Fe[1] + Pe[1];
  for (int s = 0; s < 4; ++s) free(kk[s]);
  lws_free(&w);
}

int ora_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void ora_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}
