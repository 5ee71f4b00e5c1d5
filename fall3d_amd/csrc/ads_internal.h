// ads_internal.h -- shared declarations of the gfx950 ADS kernels and the
// C-ABI host code (libfall3d_ads_hip.so). Not a public header.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

namespace f3d {

// Device copies of the wind/diffusion/settling fields, FP32, every array
// x-fastest so that y and z sweeps read them coalesced across x:
//   u1,u2 (0:nx+2, ny, nz)       idx = f + (nx+3)*((j-1) + ny*(k-1))
//   v1,v2 (nx, 0:ny+2, nz)       idx = (i-1) + nx*(f + (ny+3)*(k-1))
//   w1,w2 (nx, ny, 0:nz+2)       idx = (i-1) + nx*((j-1) + ny*f)
//   k1    (-1:nx+2, ny, nz)      idx = (i+1) + (nx+4)*((j-1) + ny*(k-1))
//   k2    (nx, -1:ny+2, nz)      idx = (i-1) + nx*((j+1) + (ny+4)*(k-1))
//   k3    (nx, ny, -1:nz+2)      idx = (i-1) + nx*((j-1) + ny*(k+1))
//   vs    (nx, ny, 0:nz+2, nb)   idx = (i-1) + nx*((j-1) + ny*(f + (nz+3)*b))
// The concentration keeps the reference's index order c(-1:nx+2,-1:ny+2,-1:nz+2,nb).
// The FP64 state stores it densely (x pitch nx+4). The FP32 state (and its
// compensated low part) pads every x row to state_pitch(nx) floats, a multiple
// of 32, and the state pointer sits 30 floats past a 256-byte-aligned
// allocation: interior cell 1 of every row starts a 128-byte line, so a
// wave's 256-byte row piece covers two whole lines.
//   chi (-1:nx+2, -1:ny+2, -1:nz+2, nb)  idx = (i+1) + P*((j+1) + (ny+4)*((k+1) + (nz+4)*b)),
//   P = state_pitch(nx); the pad words of each row stay zero.
__host__ __device__ constexpr int state_pitch(int nx) { return (nx + 4 + 31) & ~31; }
constexpr int kStateLead = 30;  // floats between the allocation and the state pointer

// Workgroup order: consecutive workgroups are dispatched round-robin over
// the 8 XCDs (each with its own L2). Renumber so that each XCD receives a
// contiguous range of logical blocks: neighbouring x tiles and the bin
// groups of the same columns then share an L2.
__device__ __forceinline__ int xcd_block(const int b, const int nblocks) {
  const int q = nblocks >> 3, r = nblocks & 7;
  const int xcd = b & 7, idx = b >> 3;
  return (xcd < r) ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

struct SweepArgs {
  float *chi;             // state (hi part)
  float *clo;             // low part (compensated state) or nullptr
  int nx, ny, nz;         // interior extents
  int n;                  // cells along the sweep axis
  int S;                  // segments (threads) per line
  int wpl;                // x sweep: waves per line (1: a wave holds one or more whole lines)
  int TX;                 // lines per workgroup (x-columns for y/z sweeps)
  int ntrans;             // x tiles of 64 columns (y, z sweeps)
  int ntrans2;            // bin groups (y, z sweeps; blockIdx.x = (line, tile) * groups + group)
  int gx;                 // workgroups along the lines (x sweep: block = (bx, bin group))
  int bin0, nbin;         // bins handled per workgroup (blockIdx.y selects the group)
  int nbins;              // total bins
  int tm;                 // 1 Euler, 2 RK4
  int write_ghosts;       // store the free-flow ghosts back into c
  int bc;                 // ghost rule: 0 free flow, 1 Dirichlet (bvalue 0), 2 zero-gradient ("periodic")
  float dt, stime;
  const float *vel1, *vel2;  // face winds along the sweep axis (device layout)
  const float *kd;           // diffusion at mass points along the sweep axis
  const float *vs;           // settling velocity (z sweep only)
  const float *hdxp;         // 0.5/dxp(f-1) for faces f = 0..n+2
  const float *idxb;         // 1/dxb(p) for cells p = 0..n+2
  const float *dtib;         // dt/dxb(p), p = 0..n+2 (for the launch's dt)
  const float *hdtib;        // 0.5*dt/dxb(p)
  const float *ftib;         // final-update factor: dt/(6 dxb(p)) for RK4, dt/dxb(p) for Euler
  const float *itab;         // streaming sweeps: 8 floats per pipeline iteration (see iter_table)
  const double *dxb64;       // dxb(p) in FP64, p = 0..n+2 (mass measure of the ledger)
  const double *areaA;       // FP64 transverse cell sizes for the face area (index 0..)
  const double *areaB;
  double *acum;              // ground accumulation (z sweep only), reference layout, FP64
  double *partials;          // [gridDim.x*gridDim.y][4]: ref flux beg/end, ledger beg/end
  double *rpartials;         // [gridDim.x*gridDim.y][4]: slot 0 the FP32 rounding residual, slots 1-3 zero
  const double *ftib64;      // final-update factor w dt/dxb(p) in FP64, p = 0..n+2 (FP32 state)
  float *fout;               // optional face mass flux output (x: u layout, y: v layout, per bin), or nullptr
  float fscale;              // fout += fscale * (G1 + 2 G2 + 2 G3 + G4) (RK4: dt/6; Euler: dt * G)
  int fout_add;              // 1: accumulate into fout; 0: overwrite (every interior face written)
};

// One launch over the members of an ensemble: member m owns the workgroups
// bstart[m] .. bstart[m+1]-1 and is described by args[m] (device memory).
struct SweepBatch {
  const SweepArgs *args;
  const int *bstart;  // nmem + 1 entries
  int nmem;
};

// Table of one axis for time step dt (ads_aux.hip), 3 (n+3) floats:
// [dt/dxb | 0.5 dt/dxb | w dt/dxb], w = 1/6 (RK4) or 1 (Euler, tm == 1),
// and the last one in FP64 (tab64, n+3 doubles).
hipError_t dt_over_dxb(const double *dxb64, int n3, double dt, int tm, float *tab, double *tab64, hipStream_t st);
// Per-iteration factors of the streaming sweep of an axis with n cells and
// nst stages: row r (8 floats) belongs to iteration t = r - kItabOffset and
// holds 0.5/dxp of face t-2, 0.5 dt/dxb of cells t-3 and t-5, dt/dxb of
// cell t-7 (these three are 0 for cells outside 1..n), the final factor
// of cell q = t-2 nst-1 (indices clamped to 0..n+2) and dxb(q) in FP32
// (slot 5, the weight of q's rounding residual). Iterations run from at
// least 2 - kPF - 15 (> -kItabOffset) to n + 2 nst + 2 (the row fetched
// one iteration ahead of the last).
// tab: the dt_over_dxb table of the axis; hdxp: 0.5/dxp(f-1) at face f.
constexpr int kItabOffset = 24;
__host__ __device__ inline int iter_rows(int n) { return (n > 0 ? n : 0) + 2 * 4 + 3 + kItabOffset; }
hipError_t iter_table(const float *tab, const float *hdxp, const double *dxb64, int n, int nst, float *itab,
                      hipStream_t st);

// Sweep launchers (ads_sweep_*.hip). cpt = cells per lane (x sweep) or
// bins per thread (y and z sweeps, 1 or 2).
hipError_t launch_sweep(int dir, int cpt, int limiter, bool comp, int nst, const SweepArgs &a, dim3 grid,
                        dim3 block, hipStream_t st);
hipError_t launch_sweep_batch(int dir, int cpt, int limiter, bool comp, int nst, const SweepBatch &b,
                              int nblocks, dim3 block, hipStream_t st);
bool sweep_cpt_supported(int cpt);
extern const int kSweepCpt[];
extern const int kSweepCptCount;

}  // namespace f3d
