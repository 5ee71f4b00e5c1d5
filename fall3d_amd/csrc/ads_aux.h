// ads_aux.h -- launchers of the kernels in ads_aux.hip (internal).
#pragma once

#include "ads_internal.h"

namespace f3d {

// Up to 16 reductions of sweep partials (4 doubles per block) into acc[].
// Up to 8 reductions of per-workgroup partials (4 doubles per workgroup)
// into acc[]: job j adds sum_b part[b*4 + s] to acc[dst[j][s]] for
// s < nslot[j].
struct ReduceJob {
  const double *part[8];
  int nblocks[8];
  int nslot[8];
  int dst[8][4];
  int count;
};

// One reduction of an ensemble step: dst[s] += sum_b part[b*4 + s], s < nslot.
struct BatchJob {
  const double *part;
  double *dst[4];
  int nblocks;
  int nslot;
};

// Sparse source of one ensemble member (see source_inject); the injected
// mass of workgroup b goes to part[b*4].
struct SourceMember {
  float *chi, *clo;
  const size_t *idx;
  const double *rate_vol, *meas;
  double *part;
  double dt;
  int nent;
  int pad;
};

struct CflArgs {
  int nx, ny, nz, nb, criterion;
  int fp64;  // 1: the fields below are FP64 (the FP64 state), else FP32
  double stime, cfl;
  const void *u1, *u2, *v1, *v2, *w1, *w2, *k1, *k2, *k3, *vs;
  const double *dx1p, *dx2p, *dx3p;  // FP64 dX*_p, index p+1 for p = -1..n+2
};

struct Solve1DArgs {
  double *c;          // n+4, positions ips-2..ipe+2
  const double *dxp;  // n+4
  const double *dxb;  // n+3, positions ibs-1..ibe+1
  const double *u;    // n+3
  const double *k;    // n+4
  double *flux;       // 2
  double dt;
  double bvalue[2];
  int bcode[2];
  int n, limiter;
};

// dense host rows of nx4 cells <-> FP32 state rows of pitch floats (rows = (ny+4)(nz+4)nb)
hipError_t state_in(const void *src, int kind, float *hi, float *lo, int nx4, int pitch, size_t rows, hipStream_t st);
hipError_t state_out(const float *hi, const float *lo, void *dst, int kind, int nx4, int pitch, size_t rows,
                     hipStream_t st);
hipError_t to_f32(const void *src, int kind, float *dst, size_t n, hipStream_t st);
hipError_t to_f64(const void *src, int kind, double *dst, size_t n, hipStream_t st);
hipError_t transpose_to_f32(const void *src, int kind, float *dst, int A, int B, int K, hipStream_t st);
hipError_t transpose_to_f64(const void *src, int kind, double *dst, int A, int B, int K, hipStream_t st);
// FP64 device array -> n values of the host kind (4 or 8 bytes) in device memory
hipError_t from_f64(const double *src, void *dst, int kind, size_t n, hipStream_t st);
hipError_t from_f32(const float *src, void *dst, int kind, size_t n, hipStream_t st);
hipError_t lerp(const float *a, const float *b, float *out, float s, size_t n, hipStream_t st);
hipError_t source_inject(float *chi, float *clo, const size_t *idx, const double *rate_vol, const double *meas,
                         int nent, double dt, double *partials, int *nblocks, hipStream_t st);
hipError_t finalize(double *acc, const ReduceJob &jobs, hipStream_t st);
hipError_t finalize_batch(const BatchJob *jobs, int njobs, hipStream_t st);
hipError_t source_inject_batch(const SourceMember *tab, const int *bstart, int nmem, int nblocks, hipStream_t st);
hipError_t cfl_time_step(const CflArgs &a, double *part, int part_cap, double *out, hipStream_t st);
// FP64 spacings of the two mass measures: p1..p3 the mass-point spacings dX_p
// (position p at index p+1), b1..b3 the cell sizes dX_b (index p)
struct MassSpacings {
  const double *p1, *p2, *p3;
  const double *b1, *b2, *b3;
};
// state: FP64 c64 when given, else chi (+ clo). out[0] = sum of c dXp dYp dZp
// (mass_volume), out[1] = sum of c dXb dYb dZb over the interior; part holds
// part_cap doubles of per-workgroup sums
hipError_t mass(const float *chi, const float *clo, const double *c64, int nx, int ny, int nz, int nb,
                const MassSpacings &sp, double *part, int part_cap, double *out, hipStream_t st);
hipError_t solve_1d(const Solve1DArgs &a, hipStream_t st);
// non-finite cells, negative cells, smallest finite value of the interior -> out[0..2]
hipError_t health(const float *chi, const float *clo, const double *c64, int nx, int ny, int nz, int nb,
                  double *part, int part_cap, double *out, hipStream_t st);
// n floats src -> dst with 16-byte buffer loads/stores (HBM counter calibration)
hipError_t copy_probe(const float *src, float *dst, size_t n, hipStream_t st);
hipError_t copy_probe64(const double *src, double *dst, size_t n, hipStream_t st);

}  // namespace f3d
