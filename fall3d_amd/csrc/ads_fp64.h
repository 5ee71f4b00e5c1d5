// ads_fp64.h -- launchers of the FP64-state kernels (ads_fp64.hip), internal.
#pragma once

#include "ads_internal.h"

namespace f3d {

// One sweep of the FP64 state. A thread owns a whole line (all of its n
// cells and 4 ghosts) of one bin; lines are numbered with the fastest
// transverse index first (x sweep: j then k; y sweep: i then k; z sweep:
// i then j), bins outermost. Device layouts of vel/kd/vs are those of the
// FP32 fields (ads_internal.h), in FP64, except for the x sweep, which
// reads the state, wind and diffusion with their two fastest axes swapped
// (c(-1:ny+2,-1:nx+2,-1:nz+2,nb), vel(ny,0:nx+2,nz), kd(ny,-1:nx+2,nz)), so
// that neighbouring lines (consecutive j) are adjacent in memory.
struct Sweep64Args {
  double *c;                  // c(-1:nx+2,-1:ny+2,-1:nz+2,nb) (x sweep: transposed, see above)
  int nx, ny, nz, nb;
  int n;                      // cells along the sweep axis
  int tm;                     // 1 Euler, 2 RK4
  int bc;                     // ghost rule: 0 free flow, 1 Dirichlet (bvalue 0), 2 zero-gradient
  long nlines;                // lines per bin
  double dt, stime;
  const double *vel1, *vel2;  // face winds along the axis; (1-stime) vel1 + stime vel2
  const double *kd;           // diffusion at mass points along the axis
  const double *vs;           // settling velocity (z sweep), subtracted from the wind
  const double *dxp;          // dX_p of the axis, position p at index p+1 (p = -1..n+2)
  const double *dxb;          // dX_b of the axis, position p at index p (p = 0..n+2)
  const double *areaA, *areaB;  // transverse dX_b: face area areaB(t2) areaA(t1)
  double *acum;               // ground accumulation (z sweep)
  double *partials;           // [gridDim.x][4]: reference flux beg/end, ledger beg/end
};

// workgroups of a sweep over nlines lines of nb bins (grid-stride beyond)
int sweep64_blocks(long nlines, int nb);
constexpr int kSweep64MaxBlocks = 8192;
hipError_t launch_sweep64(int dir, int limiter, const Sweep64Args &a, int nblocks, hipStream_t st);

// c[idx] += dt * s / vol per entry (Sources/mod_F3D.F90:1456-1473); the
// injected mass (increment times meas) summed into partials[b*4].
hipError_t source_inject64(double *c, const size_t *idx, const double *s, const double *vol, const double *meas,
                           int nent, double dt, double *partials, int *nblocks, hipStream_t st);

}  // namespace f3d
