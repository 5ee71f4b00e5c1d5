/*
 * sanitize_driver.c -- runs every function of the FP64 oracle on small
 * random problems so that the oracle can be built and checked under
 * AddressSanitizer and UndefinedBehaviorSanitizer (make -C oracle sanitize).
 *
 * TEST INFRASTRUCTURE, like ads_oracle.c: run only by the CPU test suite.
 * Every array is allocated to exactly the size of its reference bounds
 * (see the header of ads_oracle.c), so any access outside them is reported.
 * Exit status 0 and "sanitize_driver: ok" on success.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

void ora_solve_x(const int *cfg, const double *const *grid, double dt, double *c, const double *k1,
                 const double *u, double *Wf, double *Ef, double *ledger);
void ora_solve_y(const int *cfg, const double *const *grid, double dt, double *c, const double *k2,
                 const double *v, double *Sf, double *Nf, double *ledger);
void ora_solve_z(const int *cfg, const double *const *grid, double dt, double *c, double *acum,
                 const double *k3, const double *w1, const double *w2, double stime, const double *vs,
                 double *Df, double *Uf, double *ledger);
double ora_add_source_metric(const int *cfg, const double *const *grid, double dt, double *c, const double *s,
                             const double *hm1, const double *hm2, const double *hm3);
double ora_time_step(const int *cfg, const double *const *grid, const double *u, const double *v,
                     const double *w, const double *k1, const double *k2, const double *k3, const double *vs,
                     int criterion, double cfl);
double ora_mass_volume(const int *cfg, const double *const *grid, const double *c);
double ora_mass_cells(const int *cfg, const double *const *grid, const double *c);
double ora_step(const int *cfg, const double *const *grid, double dt, int iiter, double *c, double *acum,
                const double *u, const double *v, const double *w1, const double *w2, double stime,
                const double *k1, const double *k2, const double *k3, const double *vs, const double *s,
                double *flux6, double *ledger6);
void ora_solve_1d(int n, double *cl, const double *dxpl, const double *dxbl, const double *ul, const double *kl,
                  double dt, const double *bvalue, double *flux, const int *bcode, int lim);

static uint64_t rng_state = 0x9e3779b97f4a7c15ull;

static double urand(double lo, double hi) {
  /* SplitMix64 */
  uint64_t z = (rng_state += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  z ^= z >> 31;
  return lo + (hi - lo) * (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

static double *fill(size_t n, double lo, double hi) {
  double *a = malloc(n * sizeof(double));
  for (size_t i = 0; i < n; ++i) a[i] = urand(lo, hi);
  return a;
}

static int run_problem(int nx, int ny, int nz, int nb, int lim, int tm, double *checksum) {
  int cfg[13] = {nx, ny, nz, nb, 1, 1, 1, 1, 1, 1, lim, tm, 0};
  double *g[6];
  g[0] = fill((size_t)nx + 4, 800.0, 1200.0);
  g[1] = fill((size_t)nx + 3, 800.0, 1200.0);
  g[2] = fill((size_t)ny + 4, 800.0, 1200.0);
  g[3] = fill((size_t)ny + 3, 800.0, 1200.0);
  g[4] = fill((size_t)nz + 4, 100.0, 300.0);
  g[5] = fill((size_t)nz + 3, 100.0, 300.0);
  const size_t nc = (size_t)(nx + 4) * (ny + 4) * (nz + 4) * nb;
  double *c = fill(nc, 0.0, 1.0);
  for (size_t i = 0; i < nc; i += 7) c[i] = 0.0; /* plateaus and zeros: EPSILON branches */
  double *acum = fill((size_t)(nx + 4) * (ny + 4) * nb, 0.0, 0.0);
  double *u = fill((size_t)(nx + 3) * ny * nz, -5.0, 5.0);
  double *v = fill((size_t)(ny + 3) * nx * nz, -5.0, 5.0);
  double *w1 = fill((size_t)(nz + 3) * nx * ny, -1.0, 1.0);
  double *w2 = fill((size_t)(nz + 3) * nx * ny, -1.0, 1.0);
  double *k1 = fill((size_t)(nx + 4) * ny * nz, 10.0, 200.0);
  double *k2 = fill((size_t)(ny + 4) * nx * nz, 10.0, 200.0);
  double *k3 = fill((size_t)(nz + 4) * nx * ny, 1.0, 20.0);
  double *vs = fill((size_t)nx * ny * (nz + 3) * nb, 0.0, 1.0);
  double *s = fill((size_t)nx * ny * nz * nb, 0.0, 10.0);
  double *hm1 = fill((size_t)ny, 0.5, 1.0);
  double *hm2 = fill((size_t)ny + 4, 1.0, 1.0);
  double *hm3 = fill((size_t)nx * ny, 0.7, 1.0);
  double flux[6] = {0}, ledger[6] = {0};
  double dt = ora_time_step(cfg, (const double *const *)g, u, v, w1, k1, k2, k3, vs, 2, 0.9);
  dt += ora_time_step(cfg, (const double *const *)g, u, v, w1, k1, k2, k3, vs, 1, 0.9) * 0.0;
  double acc = 0.0;
  acc += ora_add_source_metric(cfg, (const double *const *)g, dt, c, s, hm1, hm2, hm3);
  acc += ora_add_source_metric(cfg, (const double *const *)g, dt, c, s, NULL, NULL, NULL);
  ora_solve_x(cfg, (const double *const *)g, dt, c, k1, u, &flux[0], &flux[1], &ledger[0]);
  ora_solve_y(cfg, (const double *const *)g, dt, c, k2, v, &flux[2], &flux[3], &ledger[2]);
  ora_solve_z(cfg, (const double *const *)g, dt, c, acum, k3, w1, w2, 0.3, vs, &flux[4], &flux[5], &ledger[4]);
  for (int it = 0; it < 2; ++it)
    acc += ora_step(cfg, (const double *const *)g, dt, it, c, acum, u, v, w1, w2, 0.0, k1, k2, k3, vs, s, flux,
                    ledger);
  acc += ora_mass_volume(cfg, (const double *const *)g, c) + ora_mass_cells(cfg, (const double *const *)g, c);
  for (int i = 0; i < 6; ++i) acc += flux[i] + ledger[i];
  for (size_t i = 0; i < (size_t)(nx + 4) * (ny + 4) * nb; ++i) acc += acum[i];
  /* ADS_solve_1D on every boundary code pair */
  const int n = nx;
  static const int codes[4] = {0, -1, 1, 2};
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      double *cl = fill((size_t)n + 4, 0.0, 1.0), *ul = fill((size_t)n + 3, -2.0, 2.0);
      double *kl = fill((size_t)n + 4, 5.0, 50.0);
      double bv[2] = {0.2, -0.1}, fl[2] = {0.0, 0.0};
      int bc[2] = {codes[a], codes[b]};
      ora_solve_1d(n, cl, g[0], g[1], ul, kl, 10.0, bv, fl, bc, lim);
      for (int i = 0; i < n + 4; ++i) acc += cl[i];
      acc += fl[0] + fl[1];
      free(cl);
      free(ul);
      free(kl);
    }
  for (int i = 0; i < 6; ++i) free(g[i]);
  free(c); free(acum); free(u); free(v); free(w1); free(w2); free(k1); free(k2); free(k3); free(vs); free(s);
  free(hm1); free(hm2); free(hm3);
  if (!isfinite(acc)) return 1;
  *checksum += acc;
  return 0;
}

int main(void) {
  double sum = 0.0;
  static const int shapes[3][4] = {{6, 5, 4, 2}, {3, 7, 5, 1}, {9, 2, 3, 3}};
  for (int sh = 0; sh < 3; ++sh)
    for (int lim = 1; lim <= 3; ++lim)
      for (int tm = 1; tm <= 2; ++tm)
        if (run_problem(shapes[sh][0], shapes[sh][1], shapes[sh][2], shapes[sh][3], lim, tm, &sum)) {
          fprintf(stderr, "sanitize_driver: non-finite result (shape %d, limiter %d, marching %d)\n", sh, lim, tm);
          return 1;
        }
  printf("sanitize_driver: ok (checksum %.6e)\n", sum);
  return 0;
}
