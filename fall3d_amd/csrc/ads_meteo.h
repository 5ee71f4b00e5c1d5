// ads_meteo.h -- the device meteo update (ads_meteo.hip), internal.
#pragma once

#include "ads_internal.h"

namespace f3d {

// Inputs (FP64 device copies) and outputs (the sweeps' device layouts) of
// one meteo update. Corner fields are (nx+1)(ny+1)(nz+1), x fastest, index
// (i-1) + (nx+1)((j-1) + (ny+1)(k-1)); 2-D corner fields (nx+1)(ny+1).
// [0] = the interval's first meteo time, [1] = its second.
struct MeteoArgs {
  int nx, ny, nz, nb;
  int map_h, map_v, modkh, modkv, modv;
  double kh0, kv0, rams_cs, dlon, dlat, x3max, s_time, vset_fac;
  double *uc[2], *vc[2], *wc[2];             // winds (ground corners zeroed on the device)
  const double *tc[2], *tvc[2], *rhoc[2];    // temperature, virtual temperature, air density
  const double *pblhc[2], *ustc[2], *monc[2];  // boundary-layer height, u*, Monin-Obukhov length
  const double *zc, *hc;                     // corner heights, corner topography
  const double *lusec;                       // land use at the corners (dry deposition)
  int drydep;                                // 1: add the dry deposition velocity at faces 0 and 1
  const double *hm1;                         // Hm1_p(1:ny)
  const double *hm2;                         // Hm2_p(-1:ny+2)
  const double *hm3, *dhdx, *dhdy;           // (nx, ny)
  const double *x3c, *x3p;                   // X3_c(1:nz+1), X3_p(1:nz)
  const double *dx1p, *dx2p;                 // dX1_p(-1:nx+2), dX2_p(-1:ny+2)
  const double *diam, *rhop, *psi;           // per bin
  const int *cat, *aggregate;                // per bin
  unsigned *flag;                            // bit 0: a settling velocity did not converge
  // outputs: [0] first time, [1] second time (winds); FP64 copies or nullptr
  float *u32[2], *v32[2], *w32[2], *k1_32, *k2_32, *k3_32, *vs32;
  double *u64[2], *v64[2], *w64[2], *k1_64, *k2_64, *k3_64, *vs64;
};

hipError_t meteo_update(const MeteoArgs &a, hipStream_t st);

}  // namespace f3d
