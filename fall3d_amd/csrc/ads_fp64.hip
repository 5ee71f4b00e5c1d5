// ads_fp64.hip -- the FP64 state (FALL3D_ADS_STATE_FP64): the KT sweeps of
// Sources/mod_ADS.F90 (ADS_solve_along_x/y/z :233-815, kt_rhs :1057-1120,
// flux limiters :1124-1223, free-flow ghosts :1239-1324) and the source term
// (Sources/mod_F3D.F90:1456-1473) in double precision, every statement
// rounded as the reference writes it (the library is built without
// floating-point contraction), so that the state follows an FP64 CPU
// statement of the reference bit for bit.
//
// A thread integrates one line of one bin in one pass over its cells. The
// RK4 stages run as a pipeline along the line: at step t stage 1 forms the
// flux of face t, stage s that of face t - 2(s-1). Face f of stage s needs
// stage s's input at cells f-2 .. f+1, and stage s-1 finished cell f+1 at
// its face f+2, in the same step: each value of the line is read once and
// written once, and the stage inputs and the RK4 accumulator live in
// registers (short shift registers of 7 entries).
#include <cmath>

#include "ads_fp64.h"

namespace f3d {
namespace {

constexpr double kEpsilon = 1e-12;  // EPSILON, Sources/mod_KindType.F90:42

// flux-limiter ratio r of cell i from c(i-1), c(i), c(i+1)
// (minmod 1, superbee 2, ospre 3; Sources/mod_ADS.F90:1124-1223)
template <int LIM>
__device__ __forceinline__ double limiter64(double cm, double c0, double cp) {
  const double dp = cp - c0;
  if (fabs(dp) < kEpsilon) {
    if constexpr (LIM == 1) {
      return 1.0;
    } else {
      const double dm = c0 - cm;
      if (fabs(dm) < kEpsilon) return 0.0;
      return LIM == 2 ? 2.0 : 1.5;
    }
  }
  const double rr = (c0 - cm) / (cp - c0);
  if constexpr (LIM == 1) {
    const double t = fmin(1.0, rr);
    return fmax(0.0, t);
  } else if constexpr (LIM == 2) {
    const double x = fmin(2.0 * rr, 1.0);
    const double y = fmin(rr, 2.0);
    const double m = fmax(0.0, x);
    return fmax(m, y);
  } else {
    return 1.5 * (rr * rr + rr) / (rr * rr + rr + 1.0);
  }
}

// One stage's window on its line: X(f-1), X(f), r(f-1), and the advective
// and diffusive fluxes of face f-1.
struct Win {
  double xm, x0, rm, Fm, Pm;
};

// Moves a stage to face f given X(f+1) = xp, the face wind u, the face
// diffusion kf = 0.5 (k(f) + k(f-1)) and dX_p(f-1). For f >= 1 the face's
// advective flux F and diffusive flux P are returned (face 0 only forms the
// ratio r(0) its right neighbour needs).
template <int LIM>
__device__ __forceinline__ void advance(Win &w, int f, double xp, double u, double kf, double dxpf, double &F,
                                        double &P) {
  const double r = limiter64<LIM>(w.xm, w.x0, xp);
  if (f >= 1) {
    const double cl = w.xm + 0.5 * w.rm * (w.x0 - w.xm);  // left state at face f
    const double cr = w.x0 - 0.5 * r * (xp - w.x0);       // right state at face f
    const double Fl = cl * u, Fr = cr * u;
    F = 0.5 * (Fl + Fr) - 0.5 * fabs(u) * (cr - cl);
    P = kf * (w.x0 - w.xm) / dxpf;
  }
  w.xm = w.x0;
  w.x0 = xp;
  w.rm = r;
}

__device__ __forceinline__ double wsum64(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// partials[blockIdx.x*4 + s] = workgroup sum of v[s] (fixed order)
__device__ __forceinline__ void block_store4(const double (&v)[4], double *partials) {
  __shared__ double red[4][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const double r = wsum64(v[s]);
    if (lane == 0) red[s][wid] = r;
  }
  __syncthreads();
  if (threadIdx.x < 4) {
    double sum = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) sum += red[threadIdx.x][w];
    partials[(size_t)blockIdx.x * 4 + threadIdx.x] = sum;
  }
}

template <int DIR, int LIM, int NST>
__global__ void __launch_bounds__(256) k_sweep64(Sweep64Args a) {
  constexpr int LAG = 2 * (NST - 1);  // faces between stage 1 and the last stage
  const int n = a.n, nx = a.nx, ny = a.ny, nz = a.nz;
  // Thread order: 64 consecutive lines of one bin per wave, the waves after
  // it the same lines for the other bins, and the blocks renumbered so that
  // consecutive ones share an XCD: every bin of a line reads the line's winds
  // and diffusion close together in time, through one L2.
  const long cw = 64L * a.nb;  // threads per chunk of 64 lines
  const long total = (a.nlines + 63) / 64 * cw;
  const long lb = xcd_block(blockIdx.x, gridDim.x);
  const double s = a.stime, s1 = 1.0 - a.stime, dt = a.dt;
  const bool lerp = DIR == 2 || s != 0.0;  // x/y winds at stime 0 are vel1 as given
  const double *dxp = a.dxp + 1;           // dxp[p], p = -1..n+2
  const double *dxb = a.dxb;               // dxb[p], p = 0..n+2
  double acc[4] = {0.0, 0.0, 0.0, 0.0};    // reference flux beg/end, ledger beg/end
  for (long g = lb * blockDim.x + threadIdx.x; g < total; g += (long)gridDim.x * blockDim.x) {
    const long chunk = g / cw;
    const int r = (int)(g - chunk * cw);
    const int b = r >> 6;
    const long L = chunk * 64 + (r & 63);
    if (L >= a.nlines) continue;  // past the last line (the last chunk's idle lanes)
    int t1, t2;  // transverse indices (x: j, k; y: i, k; z: i, j)
    ptrdiff_t cb, cs;
    size_t vb, vst, kb, sb = 0;
    if constexpr (DIR == 0) {
      // (x, y)-transposed layouts: c(j, p, k, b), wind(j, f, k), kd(j, p, k)
      t1 = (int)(L % ny) + 1;
      t2 = (int)(L / ny) + 1;
      cb = (ptrdiff_t)((size_t)(t1 + 1) + (size_t)(ny + 4) * (1 + (size_t)(nx + 4) * ((size_t)(t2 + 1) + (size_t)(nz + 4) * b)));
      cs = ny + 4;
      vb = (size_t)(t1 - 1) + (size_t)ny * (nx + 3) * (t2 - 1);
      kb = (size_t)(t1 - 1) + (size_t)ny * (nx + 4) * (t2 - 1);
      vst = ny;
    } else if constexpr (DIR == 1) {
      t1 = (int)(L % nx) + 1;
      t2 = (int)(L / nx) + 1;
      cb = (ptrdiff_t)((size_t)(t1 + 1) + (size_t)(nx + 4) * (1 + (size_t)(ny + 4) * ((size_t)(t2 + 1) + (size_t)(nz + 4) * b)));
      cs = nx + 4;
      vb = (size_t)(t1 - 1) + (size_t)nx * (ny + 3) * (t2 - 1);
      kb = (size_t)(t1 - 1) + (size_t)nx * (ny + 4) * (t2 - 1);
      vst = nx;
    } else {
      t1 = (int)(L % nx) + 1;
      t2 = (int)(L / nx) + 1;
      cb = (ptrdiff_t)((size_t)(t1 + 1) + (size_t)(nx + 4) * ((size_t)(t2 + 1) + (size_t)(ny + 4) * (1 + (size_t)(nz + 4) * b)));
      cs = (ptrdiff_t)(nx + 4) * (ny + 4);
      vb = (size_t)L;
      kb = (size_t)L;
      vst = (size_t)nx * ny;
      sb = (size_t)L + (size_t)nx * ny * (nz + 3) * b;
    }
    double *line = a.c + cb;  // line[p * cs] = c at position p of the line (p = -1..n+2)
    auto wind = [&](int f) -> double {
      const size_t o = vb + (size_t)f * vst;
      if (!lerp) return a.vel1[o];
      const double v = s1 * a.vel1[o] + s * a.vel2[o];
      if constexpr (DIR == 2) return v - a.vs[sb + (size_t)f * vst];
      return v;
    };
    auto kval = [&](int p) -> double { return a.kd[kb + (size_t)(p + 1) * vst]; };

    // boundary ghosts: with free flow, inflow at a side (u(1) > 0,
    // u(n+1) < 0) gives 2 bvalue - c with bvalue = 0 and outflow a
    // zero-gradient copy; Dirichlet always gives the former, the
    // zero-gradient rule always the latter; the outer ghost copies the inner
    const double uend = wind(n + 1);
    const double c1 = line[cs];
    const bool inw = a.bc == 0 ? (wind(1) > 0.0) : (a.bc == 1);
    const bool ine = a.bc == 0 ? (uend < 0.0) : (a.bc == 1);
    const double gw = inw ? 2.0 * 0.0 - c1 : c1;
    line[0] = gw;
    line[-cs] = gw;

    Win w[NST];
#pragma unroll
    for (int q = 0; q < NST; ++q) w[q] = Win{gw, gw, 0.0, 0.0, 0.0};
    // lr[j] = L(t+1-j) (before step 0: L(0) = L(-1) = the west ghost),
    // ncr[j] = nc(t-1-j), fu[j] and fk[j] = wind and kf of face t-j
    double lr[7], ncr[7], fu[7], fk[7];
#pragma unroll
    for (int j = 0; j < 7; ++j) {
      lr[j] = gw;
      ncr[j] = fu[j] = fk[j] = 0.0;
    }
    double fr0 = 0.0, fr1 = 0.0, lg0 = 0.0, lg1 = 0.0;

    for (int t = 0; t <= n + 1 + LAG; ++t) {
#pragma unroll
      for (int j = 6; j > 0; --j) {
        lr[j] = lr[j - 1];
        ncr[j] = ncr[j - 1];
        fu[j] = fu[j - 1];
        fk[j] = fk[j - 1];
      }
      // lead value L(t+1) (east ghosts formed from L(n) and stored)
      {
        const int p = t + 1;
        double v = 0.0;
        if (p == 1) {
          v = c1;
        } else if (p <= n) {
          v = line[p * cs];
        } else if (p == n + 1) {
          v = ine ? 2.0 * 0.0 - lr[1] : lr[1];
          line[p * cs] = v;
        } else if (p == n + 2) {
          v = lr[1];
          line[p * cs] = v;
        }
        lr[0] = v;
      }
      if (t >= 1 && t <= n + 1) {
        fu[0] = wind(t);
        fk[0] = 0.5 * (kval(t) + kval(t - 1));
      }
      // stages; pv carries stage s's finished cell to stage s+1 (same step)
      double pv = lr[0];
#pragma unroll
      for (int q = 0; q < NST; ++q) {
        const int f = t - 2 * q;  // face of this stage
        const int i = f - 1;      // cell completed at this face
        const double Li = lr[2 * q + 2 <= 6 ? 2 * q + 2 : 6];  // L(i) (unused by the last RK4 stage)
        double out = Li;  // stage input of cell i for the next stage (ghosts keep L)
        if (f >= 0 && f <= n + 1) {
          double F = 0.0, P = 0.0;
          advance<LIM>(w[q], f, pv, fu[2 * q], fk[2 * q], dxp[f - 1], F, P);
          if (f >= 1) {
            if (f >= 2) {
              const double rhs = -1.0 * (F - w[q].Fm) / dxb[i] + (P - w[q].Pm) / dxb[i];
              if constexpr (NST == 1) {
                line[i * cs] = Li + dt * rhs;
              } else if (q == 0) {
                out = Li + 0.5 * dt * rhs;
                ncr[0] = Li + dt * rhs / 6.0;
              } else if (q == 1) {
                out = Li + 0.5 * dt * rhs;
                ncr[2] = ncr[2] + dt * rhs / 3.0;
              } else if (q == 2) {
                out = Li + dt * rhs;
                ncr[4] = ncr[4] + dt * rhs / 3.0;
              } else {
                line[i * cs] = ncr[6] + dt * rhs / 6.0;
              }
            }
            // end-face ledger (RK-weighted F - P) and the last stage's F + P
            if (f == 1 || f == n + 1) {
              const double d = NST == 1 ? F - P : (F - P) / ((q == 0 || q == 3) ? 6.0 : 3.0);
              if (f == 1) {
                lg0 = (q == 0) ? d : lg0 + d;
                if (q == NST - 1) fr0 = F + P;
              } else {
                lg1 = (q == 0) ? d : lg1 + d;
                if (q == NST - 1) fr1 = F + P;
              }
            }
            w[q].Fm = F;
            w[q].Pm = P;
          }
        }
        pv = out;
      }
    }
    const double area = a.areaB[t2] * a.areaA[t1];
    acc[0] = acc[0] + fr0 * area * dt;
    acc[1] = acc[1] + fr1 * area * dt;
    acc[2] = acc[2] - lg0 * area * dt;  // mass leaving through the first face
    acc[3] = acc[3] + lg1 * area * dt;  // mass leaving through the last face
    if constexpr (DIR == 2) {
      double *ap = a.acum + (size_t)(t1 + 1) + (size_t)(nx + 4) * ((size_t)(t2 + 1) + (size_t)(ny + 4) * b);
      *ap = *ap + fabs(fr0) * dt;
    }
  }
  block_store4(acc, a.partials);
}

template <int DIR, int LIM>
hipError_t launch_lim(int tm, const Sweep64Args &a, int nblocks, hipStream_t st) {
  if (tm == 1)
    hipLaunchKernelGGL((k_sweep64<DIR, LIM, 1>), dim3(nblocks), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL((k_sweep64<DIR, LIM, 4>), dim3(nblocks), dim3(256), 0, st, a);
  return hipGetLastError();
}

template <int DIR>
hipError_t launch_dir(int limiter, const Sweep64Args &a, int nblocks, hipStream_t st) {
  if (limiter == 1) return launch_lim<DIR, 1>(a.tm, a, nblocks, st);
  if (limiter == 2) return launch_lim<DIR, 2>(a.tm, a, nblocks, st);
  return launch_lim<DIR, 3>(a.tm, a, nblocks, st);
}

__global__ void k_source64(double *__restrict__ c, const size_t *__restrict__ idx, const double *__restrict__ s,
                           const double *__restrict__ vol, const double *__restrict__ meas, int nent, double dt,
                           double *__restrict__ partials) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double v[4] = {0.0, 0.0, 0.0, 0.0};
  if (e < nent) {
    const size_t i = idx[e];
    const double inc = dt * s[e] / vol[e];
    c[i] = c[i] + inc;
    v[0] = inc * meas[e];
  }
  block_store4(v, partials);
}

}  // namespace

int sweep64_blocks(long nlines, int nb) {
  const long g = ((nlines + 63) / 64 * 64 * (long)nb + 255) / 256;
  return (int)(g < 1 ? 1 : (g > kSweep64MaxBlocks ? kSweep64MaxBlocks : g));
}

hipError_t launch_sweep64(int dir, int limiter, const Sweep64Args &a, int nblocks, hipStream_t st) {
  (void)hipGetLastError();  // drop errors of earlier, unrelated calls
  if (dir == 0) return launch_dir<0>(limiter, a, nblocks, st);
  if (dir == 1) return launch_dir<1>(limiter, a, nblocks, st);
  return launch_dir<2>(limiter, a, nblocks, st);
}

hipError_t source_inject64(double *c, const size_t *idx, const double *s, const double *vol, const double *meas,
                           int nent, double dt, double *partials, int *nblocks, hipStream_t st) {
  (void)hipGetLastError();
  const int nb = (nent + 255) / 256;
  *nblocks = nb;
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_source64, dim3(nb), dim3(256), 0, st, c, idx, s, vol, meas, nent, dt, partials);
  return hipGetLastError();
}

}  // namespace f3d
