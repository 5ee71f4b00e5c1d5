// ads_sweep_impl.h -- the directional-split Kurganov-Tadmor sweeps on gfx950.
//
// Sources/mod_ADS.F90:233-811 integrate, for every grid line of the sweep
// axis and every bin, one RK4 (or Euler) step of the 1-D KT scheme with the
// free-flow ghosts frozen for the whole sweep. Two kernels do that here:
//
//  k_stream  (y and z sweeps): one thread per line (one or two bins);
//    lanes of a wave hold 64 consecutive x-columns. The thread walks its
//    line once; the four RK stages form a software pipeline (stage s lags
//    stage s-1 by two cells, the width of the KT stencil), so each cell is
//    read once and written once. Register rings hold the last cells, face
//    winds and face fluxes. Rows arrive in groups of 4 (one 16-byte-per-lane
//    load per array, transposed through the wave's own LDS tile, two groups
//    in flight) and leave the same way.
//
//  k_sweep_x (x sweep): lines are contiguous in memory, so one wave owns a
//    line (or several short ones); each lane keeps C consecutive cells, the
//    stage values and the face fluxes in VGPRs and trades two edge cells
//    with its neighbours per stage through cross-lane permutes (no LDS
//    tile, no barrier). A lane reads and writes its cells in 16- and 8-byte
//    pieces. The next bin's cells are prefetched while the current bin is
//    integrated; wind and diffusion are read once per line for all bins.
//
// Numerics per face f (between cells f-1 and f), stage input y:
//   d_f = y_f - y_{f-1};  m_i = phi(d_i/d_{i+1}) d_{i+1} (limiter slope,
//         |d_{i+1}| < 1e-12 branch as Sources/mod_ADS.F90:1140-1146,1171-1176,1207-1212)
//   F_f = max(u_f,0) (y_{f-1} + m_{f-1}/2) + min(u_f,0) (y_f - m_f/2)   (:1103-1109)
//   G_f = F_f - (k_{f-1} + k_f)/(2 dxp_{f-1}) d_f                        (:1111-1112)
//   RHS_i = (G_i - G_{i+1}) / dxb_i                                       (:1114)
//   y1 = c + dt/2 RHS(c), y2 = c + dt/2 RHS(y1), y3 = c + dt RHS(y2)
//   c_new = c + dt (Gh_i - Gh_{i+1}) / dxb_i,  Gh = (G1 + 2 G2 + 2 G3 + G4)/6
// (the reference's new_c accumulation written in flux form). Boundary
// fluxes for the reference accumulators use the last stage only
// (:394-396, :788-795); the conservation ledger uses Gh.
#pragma once

#include <type_traits>

#include "ads_internal.h"

namespace f3d {

__device__ __forceinline__ size_t cidx(int i, int j, int k, int b, int nx, int ny, int nz) {
  return (size_t)(i + 1) + (size_t)state_pitch(nx) * ((size_t)(j + 1) + (size_t)(ny + 4) * ((size_t)(k + 1) + (size_t)(nz + 4) * (size_t)b));
}

// Limiter slope m = phi(d_i/d_{i+1}) * d_{i+1} (a = d_i, b = d_{i+1}).
template <int LIM>
__device__ __forceinline__ float limiter_slope(float a, float b) {
  const float eps = 1e-12f;
  const float ab = fabsf(b);
  if constexpr (LIM == 1) {  // minmod: phi*b = minmod(a, b) = med3(a, b, 0)
    const float m = __builtin_amdgcn_fmed3f(a, b, 0.0f);
    return (ab < eps) ? b : m;
  } else if constexpr (LIM == 2) {  // superbee
    const float aa = fabsf(a);
    const float mag = fmaxf(fminf(2.0f * aa, ab), fminf(aa, 2.0f * ab));
    const bool same = (a > 0.0f && b > 0.0f) || (a < 0.0f && b < 0.0f);
    const float m = same ? copysignf(mag, b) : 0.0f;
    const float e = (aa < eps) ? 0.0f : 2.0f * b;
    return (ab < eps) ? e : m;
  } else {  // ospre, unclipped rational form of the reference
    const float e = (fabsf(a) < eps) ? 0.0f : 1.5f * b;
    const float bs = (ab < eps) ? 1.0f : b;
    const float r = a / bs;
    const float q = fmaf(r, r, r);
    const float phi = 1.5f * q / (q + 1.0f);
    return (ab < eps) ? e : phi * b;
  }
}

// Compensated (hi, lo) += t: TwoSum of hi and t, error folded into lo,
// renormalised. Needs IEEE add/sub without reassociation (no fast-math).
__device__ __forceinline__ void comp_add(float &hi, float &lo, float t) {
  const float s = hi + t;
  const float bp = s - hi;
  const float e = (hi - (s - bp)) + (t - bp);
  const float l = lo + e;
  const float h2 = s + l;
  lo = l - (h2 - s);
  hi = h2;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Sum of four doubles over the workgroup into partials[block_id*4 + e].
__device__ __forceinline__ void block_partials(double v0, double v1, double v2, double v3, double *partials,
                                               size_t block_id) {
  __shared__ double red[4][16];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const double v4[4] = {v0, v1, v2, v3};
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const double r = wave_sum(v4[e]);
    if (lane == 0) red[e][wid] = r;
  }
  __syncthreads();
  if (t < 4) {
    double sum = 0.0;
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) sum += red[t][w];
    partials[block_id * 4 + t] = sum;
  }
}

// block_partials plus a fifth sum (the FP32 rounding residual) into its own
// row layout: rpartials[block_id*4] = v4, slots 1-3 zero.
__device__ __forceinline__ void block_partials5(double v0, double v1, double v2, double v3, double v4,
                                                double *partials, double *rpartials, size_t block_id) {
  __shared__ double red5[5][16];
  const int t = threadIdx.x, lane = t & 63, wid = t >> 6;
  const double v5[5] = {v0, v1, v2, v3, v4};
#pragma unroll
  for (int e = 0; e < 5; ++e) {
    const double r = wave_sum(v5[e]);
    if (lane == 0) red5[e][wid] = r;
  }
  __syncthreads();
  if (t < 8) {
    double sum = 0.0;
    if (t < 5) {
      const int nw = (blockDim.x + 63) >> 6;
      for (int w = 0; w < nw; ++w) sum += red5[t][w];
    }
    if (t < 4)
      partials[block_id * 4 + t] = sum;
    else
      rpartials[block_id * 4 + (t - 4)] = sum;
  }
}

// FP32-state final update with an FP64 rounding residual: the cell's new
// value c + F (gq - gq1) is formed in FP64 (F the FP64 final factor) and
// stored rounded to FP32; what the rounding dropped, exact in FP64, is
// scaled by the cell's weight w and summed into res. The sum runs in FP32:
// its error is ~2^-24 of the residual, itself ~2^-24 of the state, so the
// budget still closes to FP64 rounding, and FP32 keeps the accumulation off
// the FP64 pipe of the VALU-bound sweeps. Interior faces cancel between their
// two cells to FP64 rounding, whatever the spacing.
// fence: close the accumulation here. Without it the compiler defers the
// residual arithmetic of a whole unrolled block to its end and keeps every
// FP64 update alive until then (register spills); the streaming sweeps fence
// every other iteration, which leaves the scheduler room between fences.
__device__ __forceinline__ float final_update_fp64(float c, float gq, float gq1, double F, float w, bool count,
                                                   float &res, bool fence = true) {
  const double c64 = fma((double)gq - (double)gq1, F, (double)c);
  const float cn = (float)c64;
  res = fmaf((float)(c64 - (double)cn), count ? w : 0.0f, res);
  if (fence) __asm__ volatile("" : "+v"(res));
  return cn;
}

typedef __attribute__((address_space(4))) const int cint;  // scalar-loaded tables

// Member of an ensemble launch that owns workgroup bid: the last m with
// bstart[m] <= bid (uniform, scalar loads).
__device__ __forceinline__ int batch_member(const SweepBatch &b, const int bid) {
  cint *bs = (cint *)b.bstart;
  int lo = 0, hi = b.nmem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bs[mid] <= bid)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Arguments of member m, word by word through scalar loads.
__device__ __forceinline__ SweepArgs member_args(const SweepBatch &b, const int m) {
  static_assert(sizeof(SweepArgs) % 4 == 0, "SweepArgs is copied as 32-bit words");
  constexpr int nw = (int)(sizeof(SweepArgs) / 4);
  cint *src = (cint *)b.args + (size_t)m * nw;
  SweepArgs a;
  int *dst = reinterpret_cast<int *>(&a);
#pragma unroll
  for (int w = 0; w < nw; ++w) dst[w] = src[w];
  return a;
}

// =============================================================================
// k_stream: y (DIR 1) and z (DIR 2) sweeps, one thread per line
// =============================================================================

// Raw buffer access. The per-lane part of an address goes in the VGPR
// offset and the position along the line (the same for the whole wave) in
// the SGPR offset, so walking the line costs no vector address arithmetic.
// Every address the kernels form lies inside its buffer (positions are
// clamped with scalar min/max); stores that must not happen use a VGPR
// offset of kOOB, beyond every buffer, which the hardware range check drops.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
typedef __attribute__((address_space(4))) const float cfloat;  // scalar-loaded 1-D metrics
constexpr uint32_t kOOB = 0x80000000u;                          // all buffers here are < 2 GiB

__device__ __forceinline__ rsrc_t make_rsrc(const void *base, size_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float bload(rsrc_t r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
__device__ __forceinline__ void bstore(rsrc_t r, uint32_t voff, uint32_t soff, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, (int)soff, 0);
}
__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }

// prefetch distance (cells) of the streaming sweeps: rows arrive in groups
// of 4 iterations; kGD groups are in flight (fetched kGD groups ahead)
constexpr int kPF = 4;
#ifndef F3D_STREAM_GROUPS
#define F3D_STREAM_GROUPS 2
#endif
constexpr int kGD = F3D_STREAM_GROUPS;
static_assert(kGD == 1 || kGD == 2 || kGD == 4, "groups in flight must divide the 4 groups of a block");

// dynamic LDS of a streaming-sweep workgroup, per wave: one 4 x 64-float
// tile per fetched array (BPT cell arrays, w1, w2, K and BPT v_s arrays in
// z), one per bin for the outputs (stored 4 rows at once), and per bin the
// four boundary captures and two ghost values of every lane
template <int DIR, int BPT>
constexpr size_t stream_lds_bytes(unsigned threads) {
  return (size_t)(threads / 64) * ((size_t)(2 * BPT + 3 + ((DIR == 2) ? BPT : 0)) * 1024 + (size_t)BPT * 6 * 256);
}


// Calls f(integral_constant<int, U>) for U = U0 .. U1-1 (compile-time unrolled).
template <int U0, int U1, class F>
__device__ __forceinline__ void unroll_range(F &&f) {
  if constexpr (U0 < U1) {
    f(std::integral_constant<int, U0>{});
    unroll_range<U0 + 1, U1>(f);
  }
}

template <int DIR, int LIM, bool COMP, int NST, int BPT, bool FOUT = false>
__device__ __forceinline__ void stream_body(const SweepArgs &a, const int bid) {
  // the wave index is uniform across the wave: say so, so that the per-wave
  // buffer descriptors live in SGPRs
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nx = a.nx, ny = a.ny, nz = a.nz, n = a.n;
  // block -> (bin group, x tile, transverse line index); wave -> BPT bins
  const int groups = a.ntrans2;
  const int g = bid % groups;
  const int rest = bid / groups;
  const int tile = rest % a.ntrans;
  const int t2 = rest / a.ntrans + 1;  // k (y sweep) or j (z sweep)
  const int i = tile * 64 + lane + 1;
  const int ia = (i <= nx) ? i : nx;  // idle lanes address a valid column
  int bins[BPT];
  bool bok[BPT];
#pragma unroll
  for (int j = 0; j < BPT; ++j) {
    const int loc = wave * BPT + j;  // bin within the group
    bok[j] = (loc < a.nbin) && (g * a.nbin + loc < a.nbins);
    bins[j] = bok[j] ? g * a.nbin + loc : 0;
  }

  const size_t cbin = (size_t)state_pitch(nx) * (ny + 4) * (nz + 4);
  const size_t nvel = (DIR == 1) ? (size_t)nx * (ny + 3) * nz : (size_t)nx * ny * (nz + 3);
  const size_t nkd = (DIR == 1) ? (size_t)nx * (ny + 4) * nz : (size_t)nx * ny * (nz + 4);
  const size_t nvs = (size_t)nx * ny * (nz + 3);
  rsrc_t rc[BPT], rl[BPT], rvs[BPT], rf[BPT];
#pragma unroll
  for (int j = 0; j < BPT; ++j) {
    if constexpr (FOUT) rf[j] = make_rsrc(a.fout + nvel * bins[j], nvel * 4);
    rc[j] = make_rsrc(a.chi + cbin * bins[j], cbin * 4);
    rl[j] = make_rsrc(COMP ? a.clo + cbin * bins[j] : a.chi + cbin * bins[j], cbin * 4);
    rvs[j] = make_rsrc(DIR == 2 ? a.vs + nvs * bins[j] : a.vel1, (DIR == 2 ? nvs : nvel) * 4);
  }
  const rsrc_t rv1 = make_rsrc(a.vel1, nvel * 4);
  const rsrc_t rv2 = make_rsrc(a.vel2, nvel * 4);
  const rsrc_t rk = make_rsrc(a.kd, nkd * 4);
  // per-iteration factors, one 8-float row per iteration (scalar loads,
  // fetched one iteration ahead)
  typedef float f8 __attribute__((ext_vector_type(8)));
  typedef __attribute__((address_space(4))) const f8 cf8;
  cf8 *itab = (cf8 *)a.itab;
  f8 rnext;

  // lane byte offsets of position -1 (c), face 0 (winds), cell -1 (K); byte strides
  uint32_t vc, cs4, vv, vst4, vk, kst4;
  if constexpr (DIR == 1) {
    vc = (uint32_t)(4 * ((size_t)(ia + 1) + (size_t)state_pitch(nx) * ((size_t)(ny + 4) * (t2 + 1))));
    cs4 = (uint32_t)(4 * state_pitch(nx));
    vv = (uint32_t)(4 * ((size_t)(ia - 1) + (size_t)nx * (ny + 3) * (t2 - 1)));
    vst4 = (uint32_t)(4 * nx);
    vk = (uint32_t)(4 * ((size_t)(ia - 1) + (size_t)nx * (ny + 4) * (t2 - 1)));
    kst4 = (uint32_t)(4 * nx);
  } else {
    vc = (uint32_t)(4 * ((size_t)(ia + 1) + (size_t)state_pitch(nx) * (t2 + 1)));
    cs4 = (uint32_t)(4 * (size_t)state_pitch(nx) * (ny + 4));
    vv = (uint32_t)(4 * ((size_t)(ia - 1) + (size_t)nx * (t2 - 1)));
    vst4 = (uint32_t)(4 * (size_t)nx * ny);
    vk = vv;
    kst4 = vst4;
  }
  // grouped row fetches: lane L loads the 4 columns 4(L%16)..+3 of the
  // wave's tile from row L/16 of a 4-row group (lanes past nx read column 1)
  const int lq = lane & 15, lg = lane >> 4;
  const int col4 = tile * 64 + 1 + 4 * lq;
  const int c4 = (col4 <= nx) ? col4 : 1;
  uint32_t xc, xv, xk;
  if constexpr (DIR == 1) {
    xc = (uint32_t)(4 * ((size_t)(c4 + 1) + (size_t)state_pitch(nx) * ((size_t)(ny + 4) * (t2 + 1))));
    xv = (uint32_t)(4 * ((size_t)(c4 - 1) + (size_t)nx * (ny + 3) * (t2 - 1)));
    xk = (uint32_t)(4 * ((size_t)(c4 - 1) + (size_t)nx * (ny + 4) * (t2 - 1)));
  } else {
    xc = (uint32_t)(4 * ((size_t)(c4 + 1) + (size_t)state_pitch(nx) * (t2 + 1)));
    xv = (uint32_t)(4 * ((size_t)(c4 - 1) + (size_t)nx * (t2 - 1)));
    xk = xv;
  }
  const uint32_t xcl = xc + (uint32_t)lg * cs4, xvl = xv + (uint32_t)lg * vst4, xkl = xk + (uint32_t)lg * kst4;
  const bool lane_ok = (i <= nx);
  uint32_t vstore[BPT];
#pragma unroll
  for (int j = 0; j < BPT; ++j) vstore[j] = (lane_ok && bok[j]) ? vc : kOOB;
  const double area = lane_ok ? a.areaB[t2] * a.areaA[ia] : 0.0;  // dX3_b(k) dX1_b(i) / dX2_b(j) dX1_b(i)
  const float s = a.stime, s1 = 1.0f - a.stime;

  auto cpos = [&](int P) -> uint32_t { return (uint32_t)clampi(P + 1, 0, n + 3) * cs4; };   // cell P of c
  auto fpos = [&](int f) -> uint32_t { return (uint32_t)clampi(f, 0, n + 2) * vst4; };      // face f
  auto kpos = [&](int P) -> uint32_t { return (uint32_t)clampi(P + 1, 0, n + 3) * kst4; };  // cell P of K

  // free-flow ghosts from cells 1, n and faces 1, n+1 (Sources/mod_ADS.F90:1239-1324)
  float gL[BPT], gR[BPT];
  {
    const float wb = s1 * bload(rv1, vv, fpos(1)) + s * bload(rv2, vv, fpos(1));
    const float we = s1 * bload(rv1, vv, fpos(n + 1)) + s * bload(rv2, vv, fpos(n + 1));
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
      float vfirst = wb, vlast = we;
      if constexpr (DIR == 2) {
        vfirst -= bload(rvs[j], vv, fpos(1));
        vlast -= bload(rvs[j], vv, fpos(n + 1));
      }
      const float c1 = bload(rc[j], vc, cpos(1));
      const float cnn = bload(rc[j], vc, cpos(n));
      const bool inL = a.bc == 0 ? (vfirst > 0.0f) : (a.bc == 1);
      const bool inR = a.bc == 0 ? (vlast < 0.0f) : (a.bc == 1);
      gL[j] = inL ? (2.0f * 0.0f - c1) : c1;
      gR[j] = inR ? (2.0f * 0.0f - cnn) : cnn;
      const uint32_t vg = vstore[j];  // always written: the sweep reads them back
      bstore(rc[j], vg, cpos(-1), gL[j]);
      bstore(rc[j], vg, cpos(0), gL[j]);
      bstore(rc[j], vg, cpos(n + 1), gR[j]);
      bstore(rc[j], vg, cpos(n + 2), gR[j]);
      if constexpr (COMP) {
        bstore(rl[j], vg, cpos(-1), 0.0f);
        bstore(rl[j], vg, cpos(0), 0.0f);
        bstore(rl[j], vg, cpos(n + 1), 0.0f);
        bstore(rl[j], vg, cpos(n + 2), 0.0f);
      }
    }
  }

  float res[BPT];  // FP32 state: rounding residual of this lane's final updates, per bin
#pragma unroll
  for (int j = 0; j < BPT; ++j) res[j] = 0.0f;
  // rings: cells (16) and low words (8, fetched kPF ahead of the final
  // update) per bin; raw face data (8) and face coefficients (8). The wind
  // lerp and K are shared by the bins of a thread; in the z sweep the
  // settling velocity and hence the upwind coefficients are per bin.
  constexpr int BV = (DIR == 2) ? BPT : 1;
  float cn[BPT][16], lo[BPT][8], gh[BPT][8];
  float rw1[8], rw2[8], rkk[8], kf[8];
  float rws[BV][8], up[BV][8], um[BV][8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    rw1[q] = rw2[q] = rkk[q] = kf[q] = 0.0f;
#pragma unroll
    for (int j = 0; j < BV; ++j) rws[j][q] = up[j][q] = um[j][q] = 0.0f;
#pragma unroll
    for (int j = 0; j < BPT; ++j) gh[j][q] = lo[j][q] = 0.0f;
  }
  // per-stage pipeline state: previous input, last difference, left state
  // of the next face (previous input plus half its slope), last face flux.
  // Every stage starts as if it had consumed an unbounded run of west
  // ghosts (value gL: zero difference and slope).
  float sa1[BPT][NST], sd[BPT][NST], scl[BPT][NST], sG[BPT][NST];
#pragma unroll
  for (int j = 0; j < BPT; ++j)
#pragma unroll
    for (int st = 0; st < NST; ++st) {
      sa1[j][st] = scl[j][st] = gL[j];
      sd[j][st] = sG[j][st] = 0.0f;
    }
  // Row groups. Every 4 iterations a wave fetches the next 4 rows of each
  // array with one 16-byte-per-lane load (a quarter of the texture-unit
  // work of 4 single-float loads) and transposes them through its own LDS
  // tile (no barrier: one wave writes and reads it, in order). Arrays: c of
  // each bin (cells), w1, w2 (faces), K (cells), v_s of each bin (z, faces).
  // Row offsets sit in the VGPR offset, which the buffer range check covers.
  //
  // Ghosts come from the data, with no per-stage selects:
  //  - cell ring slots start as the west ghost; a fetched cell at a
  //    position below 1 becomes the west ghost, one past n the east ghost
  //    (a select only in the groups that touch a line end);
  //  - the stage factors of the iteration table are 0 outside 1..n, so a
  //    stage's value at a ghost position stays the ring value: the frozen
  //    ghosts of the reference.
  constexpr int NAC = BPT, NAW = 3 + ((DIR == 2) ? BPT : 0);  // cell arrays, face arrays (w1, w2, K, v_s)
  constexpr int NA = NAC + NAW;
  extern __shared__ float4 stream_lds[];
  // per wave: NA + BPT tiles of 4 x 64 floats, then 6 floats per bin and lane
  constexpr int kWaveFloats = (NA + BPT) * 256 + 6 * BPT * 64;
  static_assert(kWaveFloats * 4 == (int)stream_lds_bytes<DIR, BPT>(64), "LDS layout and launch size disagree");
  float4 *tile_lds = (float4 *)((float *)stream_lds + (size_t)wave * kWaveFloats);
  const float *tile_f = (const float *)tile_lds;
  float *out_f = (float *)(tile_lds + NA * 64);  // output tiles: bin j, row r, column lane
  // output rows of a group are stored by lane L for columns 4(L%16)..+3 of
  // row L/16: whole 16-byte pieces inside 1..nx, single floats where the
  // piece crosses nx (then only columns up to nx)
  bool sfull[BPT], spart[BPT];
#pragma unroll
  for (int j = 0; j < BPT; ++j) {
    sfull[j] = bok[j] && (col4 + 3 <= nx);
    spart[j] = bok[j] && (col4 <= nx) && (col4 + 3 > nx);
  }
  float4 fc[kGD][NAC], fw[kGD][NAW];  // rows in flight, one set per group ahead
#pragma unroll
  for (int k = 0; k < kGD; ++k) {
#pragma unroll
    for (int q = 0; q < NAC; ++q) fc[k][q] = float4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int q = 0; q < NAW; ++q) fw[k][q] = float4{0.0f, 0.0f, 0.0f, 0.0f};
  }
  // fetch into set k the rows of the group that starts at iteration tg + 4:
  // cells tg+3..tg+6, faces tg+2..tg+5, K of cells tg+2..tg+5
  auto fetch = [&](const int tg, auto K) {
    constexpr int k = decltype(K)::value;
    // uniform row offset (scalar multiply) plus the lane's constant part;
    // rows outside the line wrap or run past the array and come back as
    // other lines' values or 0 from the range check: only ghost positions
    // and faces outside 1..n+1 see them
    const uint32_t oc = xcl + (uint32_t)((tg + 4) * (int)cs4);
    const uint32_t ov = xvl + (uint32_t)((tg + 2) * (int)vst4);
    const uint32_t ok = xkl + (uint32_t)((tg + 3) * (int)kst4);
#pragma unroll
    for (int j = 0; j < BPT; ++j)
      fc[k][j] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rc[j], (int)oc, 0, 0));
    fw[k][0] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rv1, (int)ov, 0, 0));
    fw[k][1] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rv2, (int)ov, 0, 0));
    fw[k][2] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rk, (int)ok, 0, 0));
    if constexpr (DIR == 2)
#pragma unroll
      for (int j = 0; j < BPT; ++j)
        fw[k][3 + j] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(rvs[j], (int)ov, 0, 0));
  };
#pragma unroll
  for (int j = 0; j < BPT; ++j)
#pragma unroll
    for (int q = 0; q < 16; ++q) cn[j][q] = gL[j];

  // boundary captures (RFb, tl1, RFe, trn of each bin) live in LDS until the
  // end: each is written once per line; so do the ghost values (read only
  // in the groups at the line ends)
  float *cap_f = out_f + BPT * 4 * 64;  // [4][BPT][64]
  float *ghost_f = cap_f + 4 * BPT * 64;  // [2][BPT][64]
#pragma unroll
  for (int j = 0; j < BPT; ++j) {
    ghost_f[(0 * BPT + j) * 64 + lane] = gL[j];
    ghost_f[(1 * BPT + j) * 64 + lane] = gR[j];
  }
  const int tend = n + 2 * NST + 1;  // iteration of the last final output (cell n)
  // Whole blocks of 16 iterations that end at tend and start at or before
  // 2 - kPF, the iteration that fetches cell 1 (earlier iterations only see
  // ghosts and drop their stores).
  const int nblk = (tend + kPF - 1 + 15) / 16;
  const int tstart = tend + 1 - 16 * nblk;
  // z lines are short (60 cells take 80 iterations from tstart): when the
  // start may move 8 iterations later and still see only ghosts, the first
  // block runs its second half only. With two groups in flight the group
  // sets keep their parity (u = 8 reads set 0, as u = 0 does).
  constexpr bool kHalf = (DIR == 2) && (kGD == 2);
  const bool half = kHalf && (tstart + 8 <= 2 - kPF);
  const int tfirst = half ? tstart + 8 : tstart;
  rnext = itab[tfirst + kItabOffset];
  // FP32 state: FP64 final factor of the cell each iteration updates
  // (q = t - 2 NST - 1, clamped into the table), one iteration ahead
  typedef __attribute__((address_space(4))) const double cdouble;
  cdouble *fin64 = (cdouble *)a.ftib64;
  double fnext = COMP ? 0.0 : fin64[clampi(tfirst - 2 * NST - 1, 0, n + 2)];
  // the first group's rows are all ghosts (the zeros above become them);
  // the rows of the next kGD-1 groups are fetched here
  unroll_range<1, kGD>([&](auto K) { fetch(tfirst + 4 * (decltype(K)::value - 1), K); });

  // One pipeline step at iteration t (u = (t - tstart) mod 16, a
  // compile-time constant: ring slots are static registers).
  auto step = [&](const int t, auto U) {
    constexpr int u = decltype(U)::value;
    // [0.5/dxp(t-2), 0.5dt/dxb(t-3), 0.5dt/dxb(t-5), dt/dxb(t-7), final(q), dxb(q)], q = t-2NST-1
    const f8 row = rnext;
    rnext = itab[t + 1 + kItabOffset];
    const double fq = fnext;
    if constexpr (!COMP) fnext = fin64[clampi(t - 2 * NST, 0, n + 2)];
    if constexpr ((u & 3) == 0) {
      // ---- group start: the rows fetched kGD groups ago (cells t-1..t+2,
      // faces and K cells t-2..t+1) into the rings, then the fetch for the
      // group kGD ahead into the same set
      constexpr int k = (u >> 2) % kGD;
#pragma unroll
      for (int q = 0; q < NAC; ++q) tile_lds[q * 64 + lane] = fc[k][q];
#pragma unroll
      for (int q = 0; q < NAW; ++q) tile_lds[(NAC + q) * 64 + lane] = fw[k][q];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int j = 0; j < BPT; ++j) cn[j][(u - 1 + r) & 15] = tile_f[(j * 4 + r) * 64 + lane];
        rw1[(u - 2 + r) & 7] = tile_f[((NAC + 0) * 4 + r) * 64 + lane];
        rw2[(u - 2 + r) & 7] = tile_f[((NAC + 1) * 4 + r) * 64 + lane];
        rkk[(u - 2 + r) & 7] = tile_f[((NAC + 2) * 4 + r) * 64 + lane];
        if constexpr (DIR == 2)
#pragma unroll
          for (int j = 0; j < BPT; ++j) rws[j][(u - 2 + r) & 7] = tile_f[((NAC + 3 + j) * 4 + r) * 64 + lane];
      }
      if (t - 1 <= 0 || t + 2 > n) {  // a line end: ghost values
        // both ghosts are read unconditionally and then selected (a select
        // between addresses would put the ring in scratch memory)
#pragma unroll
        for (int j = 0; j < BPT; ++j) {
          const float gl = ghost_f[(0 * BPT + j) * 64 + lane], gr = ghost_f[(1 * BPT + j) * 64 + lane];
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int pos = t - 1 + r;
            const float x = cn[j][(u - 1 + r) & 15];
            cn[j][(u - 1 + r) & 15] = (pos <= 0) ? gl : ((pos > n) ? gr : x);
          }
        }
      }
      fetch(t + 4 * (kGD - 1), std::integral_constant<int, k>{});
    }
    // ---- low words of the compensated state (fetched for the final update)
    if constexpr (COMP)
#pragma unroll
      for (int j = 0; j < BPT; ++j)
        lo[j][(u + kPF - 2 * NST - 1) & 7] = bload(rl[j], vc, cpos(t + kPF - 2 * NST - 1));
    // ---- coefficients of face t-2 (first used by stage 1 in this iteration)
    {
      const int sl = (u - 2) & 7;
      const float v = s1 * rw1[sl] + s * rw2[sl];
#pragma unroll
      for (int j = 0; j < BV; ++j) {
        float ve = v;
        if constexpr (DIR == 2) ve -= rws[j][sl];
        up[j][sl] = fmaxf(ve, 0.0f);
        um[j][sl] = fminf(ve, 0.0f);
      }
      kf[sl] = (rkk[(u - 3) & 7] + rkk[sl]) * row[0];
    }
    // ---- the pipeline: stage st consumes position t+1-2(st+1)
    float Glast[BPT], Plast[BPT];
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
      const int jv = (DIR == 2) ? j : 0;
      float v = cn[j][(u - 1) & 15];  // stage-0 input: position t-1
#pragma unroll
      for (int st = 0; st < NST; ++st) {
        const int fs = (u - 2 * (st + 1)) & 7;  // face t-2(st+1), completed by this stage
        const float d = v - sa1[j][st];
        const float m = limiter_slope<LIM>(sd[j][st], d);
        const float cL = scl[j][st];
        const float cR = fmaf(-0.5f, m, sa1[j][st]);
        // G = F - P with the KT flux F = max(u,0) cL + min(u,0) cR (a = |u|: upwind)
        const float Pd = kf[fs] * sd[j][st];
        const float G = fmaf(up[jv][fs], cL, fmaf(um[jv][fs], cR, -Pd));
        // RK4 weights 1 2 2 1 (the 1/6 is in ftib)
        if (st == 0)
          gh[j][fs] = G;
        else if (st == NST - 1)
          gh[j][fs] = gh[j][fs] + G;
        else
          gh[j][fs] = fmaf(2.0f, G, gh[j][fs]);
        if (st == NST - 1) {
          Glast[j] = G;
          Plast[j] = Pd;
        }
        const float Gp = sG[j][st];
        scl[j][st] = fmaf(0.5f, m, sa1[j][st]);
        sa1[j][st] = v;
        sd[j][st] = d;
        sG[j][st] = G;
        if (st < NST - 1) {
          // next stage input (position t-2st-3); the factor (0.5 dt/dxb for
          // stages 1, 2, dt/dxb for stage 3) is 0 at ghost positions
          const float hib = row[1 + st];
          v = fmaf(hib, Gp - G, cn[j][(u - 2 * (st + 1) - 1) & 15]);
        }
      }
      // ---- final update of cell q = t - 2 NST - 1 (stores before cell 1 dropped)
      {
        const int q = t - 2 * NST - 1;
        const int cl = (u - 2 * NST - 1) & 15;
        const float dtq = row[4];
        const float gq = gh[j][(u - 2 * NST - 1) & 7], gq1 = gh[j][(u - 2 * NST) & 7];
        const uint32_t vo = (q >= 1) ? vstore[j] : kOOB;
        if constexpr (COMP) {
          float hv = cn[j][cl], lv = lo[j][(u - 2 * NST - 1) & 7];
          comp_add(hv, lv, dtq * gq);
          comp_add(hv, lv, -(dtq * gq1));
          bstore(rc[j], vo, cpos(q), hv);
          bstore(rl[j], vo, cpos(q), lv);
        } else {
          (void)vo;
          out_f[(j * 4 + (u & 3)) * 64 + lane] = final_update_fp64(cn[j][cl], gq, gq1, fq, row[5], q >= 1, res[j], (u & 1) == 1);
        }
        if constexpr (FOUT) {  // face mass flux per unit area: faces q (1..n) and n+1
          const uint32_t fo = fpos(q), fo1 = fpos(q + 1);
          const uint32_t vf = (q >= 1 && vstore[j] != kOOB) ? vv : kOOB;
          const bool add = a.fout_add != 0;
          bstore(rf[j], vf, fo, fmaf(a.fscale, gq, add ? bload(rf[j], vv, fo) : 0.0f));
          const uint32_t ve = (q == n) ? vf : kOOB;
          bstore(rf[j], ve, fo1, fmaf(a.fscale, gq1, add ? bload(rf[j], vv, fo1) : 0.0f));
        }
      }
    }
    // ---- group end: the 4 output rows of each bin, cells t-2NST-4 .. t-2NST-1
    // (rows before cell 1 dropped)
    if constexpr (!COMP && (u & 3) == 3) {
      const int q0 = t - 3 - 2 * NST - 1;  // first output row of the group
      const uint32_t os = xcl + (uint32_t)((q0 + 1) * (int)cs4);
      const bool rok = (q0 + lg >= 1);
#pragma unroll
      for (int j = 0; j < BPT; ++j) {
        const float4 v = ((const float4 *)out_f)[j * 64 + lane];
        if (sfull[j] || !spart[j]) {
          const uint32_t vo = (sfull[j] && rok) ? os : kOOB;
          __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) uint32_t, v),
                                                 rc[j], (int)vo, 0, 0);
        } else {
          const int nv = nx - col4 + 1;  // 1..3 valid columns
          const float ve[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
          for (int e = 0; e < 3; ++e)
            if (e < nv && rok) bstore(rc[j], os + 4 * e, 0, ve[e]);
        }
      }
    }
    // ---- boundary captures (three iterations per line; a branch, not selects):
    // the reference's end-face fluxes F + P of the last stage at faces 1 and
    // n+1, and the final-update fluxes of cells 1 and n
    if (t == 2 * NST + 1 || t == 2 * NST + 2 || t == tend) {
      __asm__ volatile("");
      const float dtq = row[4];
#pragma unroll
      for (int j = 0; j < BPT; ++j) {
        const float rf = fmaf(2.0f, Plast[j], Glast[j]);
        const float gq = gh[j][(u - 2 * NST - 1) & 7], gq1 = gh[j][(u - 2 * NST) & 7];
        // transfers through faces 1 and n+1: FP32 products for the compensated
        // state, the raw face values for the FP32 state (scaled in FP64 below)
        if (t == 2 * NST + 1) cap_f[(0 * BPT + j) * 64 + lane] = rf;
        if (t == 2 * NST + 2) cap_f[(1 * BPT + j) * 64 + lane] = COMP ? dtq * gq : gq;
        if (t == tend) {
          cap_f[(2 * BPT + j) * 64 + lane] = rf;
          cap_f[(3 * BPT + j) * 64 + lane] = COMP ? dtq * gq1 : gq1;
        }
      }
    }
    // keep the scheduler from hoisting the scalar loads and offsets of all
    // 16 unrolled iterations to the block top (SGPR spills to VGPR lanes)
    __builtin_amdgcn_sched_barrier(0);
  };

  int t0 = tstart;
  if constexpr (kHalf) {
    if (half) {
      unroll_range<8, 16>([&](auto U) { step(t0 + decltype(U)::value, U); });
      t0 += 16;
    }
  }
  for (; t0 <= tend; t0 += 16)
    unroll_range<0, 16>([&](auto U) { step(t0 + decltype(U)::value, U); });

  double rb = 0.0, re = 0.0, lb = 0.0, le = 0.0, rr = 0.0;
  if (lane_ok) {
    const double dt = (double)a.dt;
    // FP32 state: the boundary transfers with the same FP64 factors as the
    // final updates of cells 1 and n
    const double f1 = COMP ? 1.0 : a.ftib64[1], fn = COMP ? 1.0 : a.ftib64[n];
#pragma unroll
    for (int j = 0; j < BPT; ++j) {
      if (!bok[j]) continue;
      const float RFb = cap_f[(0 * BPT + j) * 64 + lane], tl1 = cap_f[(1 * BPT + j) * 64 + lane];
      const float RFe = cap_f[(2 * BPT + j) * 64 + lane], trn = cap_f[(3 * BPT + j) * 64 + lane];
      rb += (double)RFb * area * dt;
      re += (double)RFe * area * dt;
      lb -= (double)tl1 * f1 * a.dxb64[1] * area;
      le += (double)trn * fn * a.dxb64[n] * area;
      rr += (double)res[j] * area;
      if constexpr (DIR == 2) {
        const size_t ai = (size_t)(i + 1) + (size_t)(nx + 4) * ((size_t)(t2 + 1) + (size_t)(ny + 4) * bins[j]);
        a.acum[ai] = a.acum[ai] + fabs((double)RFb) * dt;
      }
    }
  }
  block_partials5(rb, re, lb, le, rr, a.partials, a.rpartials, bid);
}

template <int DIR, int LIM, bool COMP, int NST, int BPT, bool FOUT = false>
__global__ void __launch_bounds__(512) k_stream(SweepArgs a) {
  stream_body<DIR, LIM, COMP, NST, BPT, FOUT>(a, xcd_block(blockIdx.x, gridDim.x));
}

// All members of an ensemble in one launch: the workgroup finds its member
// in the prefix table of workgroup counts, reads that member's arguments
// (scalar loads) and runs the same body on its member-local block index.
template <int DIR, int LIM, bool COMP, int NST, int BPT>
__global__ void __launch_bounds__(512) k_stream_batch(SweepBatch b) {
  const int lb = xcd_block(blockIdx.x, gridDim.x);
  const int m = batch_member(b, lb);
  const SweepArgs a = member_args(b, m);
  stream_body<DIR, LIM, COMP, NST, BPT>(a, lb - ((cint *)b.bstart)[m]);
}

// =============================================================================
// k_sweep_x: x sweep, one wave per line, C cells per lane
// =============================================================================

// C consecutive floats of one lane in as few wide accesses as possible
// (4-, 2- and 1-float pieces). The texture unit spends about the same
// cycles on a lane's 16-byte access as on a 4-byte one, so wide accesses
// cut its load per byte (addresses need only 4-byte alignment here).
typedef float f4a __attribute__((ext_vector_type(4), aligned(4)));
typedef float f2a __attribute__((ext_vector_type(2), aligned(4)));
template <int C, int Q = 0>
__device__ __forceinline__ void load_cells(const float *p, float (&v)[C]) {
  if constexpr (C - Q >= 4) {
    const f4a t = *(const f4a *)(p + Q);
    v[Q] = t.x;
    v[Q + 1] = t.y;
    v[Q + 2] = t.z;
    v[Q + 3] = t.w;
    load_cells<C, Q + 4>(p, v);
  } else if constexpr (C - Q >= 2) {
    const f2a t = *(const f2a *)(p + Q);
    v[Q] = t.x;
    v[Q + 1] = t.y;
    load_cells<C, Q + 2>(p, v);
  } else if constexpr (C - Q == 1) {
    v[Q] = p[Q];
  }
}
template <int C, int Q = 0>
__device__ __forceinline__ void store_cells(float *p, const float (&v)[C]) {
  if constexpr (C - Q >= 4) {
    *(f4a *)(p + Q) = f4a{v[Q], v[Q + 1], v[Q + 2], v[Q + 3]};
    store_cells<C, Q + 4>(p, v);
  } else if constexpr (C - Q >= 2) {
    *(f2a *)(p + Q) = f2a{v[Q], v[Q + 1]};
    store_cells<C, Q + 2>(p, v);
  } else if constexpr (C - Q == 1) {
    p[Q] = v[Q];
  }
}

template <int C, int LIM, bool COMP, int NST, bool FOUT = false>
__device__ __forceinline__ void sweep_x_body(const SweepArgs &a, const int bx, const int by) {
  static_assert(C >= 2, "segments need two edge cells");
  // Segmentation: S = n/C + 1 segments of C cells, so that face n+1 always
  // lies inside a segment (faces 0..C-1 of each lane are computed locally).
  // S <= 64: a wave holds 64/S whole lines. S > 64: wpl waves of the
  // workgroup hold one line, segment = 64 (wave in line) + lane; the lanes
  // at wave edges trade their cells and fluxes through LDS, with a
  // workgroup barrier per exchange (wpl is uniform, so the branches are).
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int S = a.S, n = a.n, wpl = a.wpl;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  int seg, line, wil = 0;
  bool act;
  if (wpl == 1) {
    const int lpw = 64 / S;  // lines per wave
    const int lslot = lane / S;
    seg = lane % S;
    line = (bx * (blockDim.x >> 6) + wave) * lpw + lslot;
    act = (lslot < lpw) && (line < ny * nz);
  } else {
    wil = wave % wpl;
    seg = wil * 64 + lane;
    line = bx * ((blockDim.x >> 6) / wpl) + wave / wpl;
    act = (seg < S) && (line < ny * nz);
  }
  const int j = act ? line % ny + 1 : 1;
  const int k = act ? line / ny + 1 : 1;
  // idle segments of a multi-wave line read line (1, 1) from its start
  const int s0 = (act || wpl == 1) ? seg * C + 1 : 1;
  // edge exchange between the waves of a line: per wave, the left edge's
  // y[C-2], y[C-1] (slots 0, 1), the right edge's y[0], G0 and gh[0]
  // (slots 2, 3, 4) and the left edge's unmodified last cell (slot 5); each
  // slot is written once per exchange phase, and a phase's reads finish
  // before the next phase's barrier
  __shared__ float xch[16][6];
  const bool xl = wpl > 1 && lane == 0 && wil > 0;         // takes from the wave to its left
  const bool xr = wpl > 1 && lane == 63 && wil < wpl - 1;  // takes from the wave to its right
  const bool first = (seg == 0) || !act;
  const bool last = (seg == S - 1) || !act;
  const int qn = n - s0;  // local index of cell n in the last segment
  const int lsrc = (lane + 63) & 63, rsrc = (lane + 1) & 63;

  const size_t cbase = cidx(0, j, k, 0, nx, ny, nz);
  const size_t cbin = (size_t)state_pitch(nx) * (ny + 4) * (nz + 4);
  const size_t vbase = (size_t)(nx + 3) * ((size_t)(j - 1) + (size_t)ny * (k - 1));
  const size_t kbase = (size_t)(nx + 4) * ((size_t)(j - 1) + (size_t)ny * (k - 1));
  const double area = a.areaB[k] * a.areaA[j];  // dX3_b(k) dX2_b(j)
  const float s = a.stime, s1 = 1.0f - a.stime;
  const float dt = a.dt;

  // wind and diffusion of this line, read once for all bins, with wide
  // unconditional loads (idle lanes read line (1, 1)). Faces past n+1 only
  // feed ghost positions, whose stage multipliers are zero, so their values
  // (the next line's, or the arrays' zeroed padding) do not matter.
  float up[C], um[C], kf[C], hb[C], db[C], fb[C];
  double f64[C];  // FP32 state: FP64 final factors of the lane's cells
  float wq[C];    // and their residual weights dxb
  float vbeg = 0.0f, vend = 0.0f;
  {
    float w1v[C], w2v[C], kk[C + 1], hx[C];
    load_cells<C>(a.vel1 + vbase + s0, w1v);
    load_cells<C>(a.vel2 + vbase + s0, w2v);
    load_cells<C + 1>(a.kd + kbase + s0, kk);  // cells s0-1 .. s0+C-1 (cell p stored at index p+1)
    load_cells<C>(a.hdxp + s0, hx);
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const float v = s1 * w1v[q] + s * w2v[q];
      up[q] = fmaxf(v, 0.0f);
      um[q] = fminf(v, 0.0f);
      kf[q] = (kk[q] + kk[q + 1]) * hx[q];
    }
    // stage multipliers are zero beyond cell n: the ghost positions of the
    // last segment then keep their frozen value through every stage
    load_cells<C>(a.dtib + s0, db);
    load_cells<C>(a.hdtib + s0, hb);
    load_cells<C>(a.ftib + s0, fb);
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const bool in = act && s0 + q <= n;
      db[q] = in ? db[q] : 0.0f;
      hb[q] = in ? hb[q] : 0.0f;
      fb[q] = in ? fb[q] : 0.0f;
    }
    if constexpr (!COMP) {
      // FP64 final factors (0 beyond cell n and on idle lanes: those
      // updates then keep the value exactly) and the cells' residual weights
#pragma unroll
      for (int q = 0; q < C; ++q) {
        const bool in = act && s0 + q <= n;
        f64[q] = in ? a.ftib64[s0 + q] : 0.0;
        wq[q] = (float)a.dxb64[s0 + q];
      }
    }

    if (act) {
      vbeg = s1 * a.vel1[vbase + 1] + s * a.vel2[vbase + 1];
      vend = s1 * a.vel1[vbase + (size_t)(n + 1)] + s * a.vel2[vbase + (size_t)(n + 1)];
    }
  }

  double acc_rb = 0.0, acc_re = 0.0, acc_lb = 0.0, acc_le = 0.0, acc_res = 0.0;
  const int bfirst = a.bin0 + by * a.nbin;
  const int blast = min(bfirst + a.nbin, a.nbins);

  // prefetch the first bin. Loads are unconditional and wide: a lane's C
  // cells may run past n (the ghosts and the next row's first cells, still
  // inside the array; replaced by the east ghost below), and idle lanes read
  // line (1, 1).
  float nxt[C], nlo[C];
  const int bpre = (bfirst < blast) ? bfirst : 0;
  load_cells<C>(a.chi + cbase + cbin * bpre + s0, nxt);
  if constexpr (COMP) load_cells<C>(a.clo + cbase + cbin * bpre + s0, nlo);

  for (int b = bfirst; b < blast; ++b) {
    const size_t cb = cbase + cbin * (size_t)b;
    float cn[C], lo[C];
#pragma unroll
    for (int q = 0; q < C; ++q) {
      cn[q] = nxt[q];
      lo[q] = nlo[q];
    }
    // prefetch the next bin while this one is integrated
    if (b + 1 < blast) {
      load_cells<C>(a.chi + cb + cbin + s0, nxt);
      if constexpr (COMP) load_cells<C>(a.clo + cb + cbin + s0, nlo);
    }
    // free-flow ghosts (Sources/mod_ADS.F90:1239-1265), frozen for the sweep.
    // Cell n is in the last segment, or (when that segment starts at n+1)
    // the last cell of the segment to its left.
    float gL = 0.0f, gR = 0.0f;
    float cleft = __shfl(cn[C - 1], lsrc, 64);
    if (wpl > 1) {
      if (lane == 63) xch[wave][5] = cn[C - 1];
      __syncthreads();
      if (xl) cleft = xch[wave - 1][5];
    }
    if (first) gL = (a.bc == 0 ? (vbeg > 0.0f) : (a.bc == 1)) ? (2.0f * 0.0f - cn[0]) : cn[0];
    if (last) {
      float cl = (qn < 0) ? cleft : cn[0];
#pragma unroll
      for (int q = 1; q < C; ++q) cl = (q == qn) ? cn[q] : cl;
      gR = (a.bc == 0 ? (vend < 0.0f) : (a.bc == 1)) ? (2.0f * 0.0f - cl) : cl;
#pragma unroll
      for (int q = 0; q < C; ++q) {
        cn[q] = (q > qn) ? gR : cn[q];
        lo[q] = (q > qn) ? 0.0f : lo[q];  // ghost low words are 0
      }
    }
    if (act && a.write_ghosts) {
      if (seg == 0) {
        a.chi[cb - 1] = gL;
        a.chi[cb] = gL;
        if constexpr (COMP) {
          a.clo[cb - 1] = 0.0f;
          a.clo[cb] = 0.0f;
        }
      }
      if (seg == S - 1) {
        a.chi[cb + (size_t)(n + 1)] = gR;
        a.chi[cb + (size_t)(n + 2)] = gR;
        if constexpr (COMP) {
          a.clo[cb + (size_t)(n + 1)] = 0.0f;
          a.clo[cb + (size_t)(n + 2)] = 0.0f;
        }
      }
    }

    float y[C], gh[C + 1];
#pragma unroll
    for (int q = 0; q < C; ++q) y[q] = cn[q];
#pragma unroll
    for (int q = 0; q <= C; ++q) gh[q] = 0.0f;
    float RFb = 0.0f, RFe = 0.0f;

#pragma unroll
    for (int st = 0; st < NST; ++st) {
      // neighbours' edge cells (same wave: cross-lane permute, no barrier)
      float yl2 = __shfl(y[C - 2], lsrc, 64);
      float yl1 = __shfl(y[C - 1], lsrc, 64);
      float yr0 = __shfl(y[0], rsrc, 64);
      if (wpl > 1) {
        if (lane == 63) {
          xch[wave][0] = y[C - 2];
          xch[wave][1] = y[C - 1];
        }
        if (lane == 0) xch[wave][2] = y[0];
        __syncthreads();
        if (xl) {
          yl2 = xch[wave - 1][0];
          yl1 = xch[wave - 1][1];
        }
        if (xr) yr0 = xch[wave + 1][2];
      }
      if (first) {
        yl2 = gL;
        yl1 = gL;
      }
      if (last) yr0 = gR;
      float w[C + 3];
      w[0] = yl2;
      w[1] = yl1;
#pragma unroll
      for (int q = 0; q < C; ++q) w[q + 2] = y[q];
      w[C + 2] = yr0;
      const bool last_stage = (st == NST - 1);
      float d_cur = w[2] - w[1];
      float m_prev = limiter_slope<LIM>(w[1] - w[0], d_cur);
      float G_prev = 0.0f, G0 = 0.0f;
#pragma unroll
      for (int q = 0; q < C; ++q) {
        const float d_next = w[q + 3] - w[q + 2];
        const float m_cur = limiter_slope<LIM>(d_cur, d_next);
        const float cL = fmaf(0.5f, m_prev, w[q + 1]);
        const float cR = fmaf(-0.5f, m_cur, w[q + 2]);
        // G = F - P with F = max(u,0) cL + min(u,0) cR
        const float P = kf[q] * d_cur;
        const float G = fmaf(up[q], cL, fmaf(um[q], cR, -P));
        if (q == 0) G0 = G;
        if (last_stage) {  // reference boundary fluxes F + P
          if (q == 0) RFb = fmaf(2.0f, P, G);
          if (q == qn + 1) RFe = fmaf(2.0f, P, G);
        }
        // RK4 weights 1 2 2 1 (the 1/6 is in fb)
        if (st == 0)
          gh[q] = G;
        else if (last_stage)
          gh[q] = gh[q] + G;
        else
          gh[q] = fmaf(2.0f, G, gh[q]);
        if (q > 0 && !last_stage) {
          y[q - 1] = fmaf((st == 2) ? db[q - 1] : hb[q - 1], G_prev - G, cn[q - 1]);
        }
        G_prev = G;
        d_cur = d_next;
        m_prev = m_cur;
      }
      // Face C is the right neighbour's face 0 (same inputs, same arithmetic:
      // bit-identical). The last segment's face C lies beyond n+1 and only
      // feeds frozen ghost positions (zero multipliers).
      float GC = __shfl(G0, rsrc, 64);
      if (wpl > 1) {
        if (lane == 0) {
          xch[wave][3] = G0;
          if (last_stage) xch[wave][4] = gh[0];
        }
        __syncthreads();
        if (xr) GC = xch[wave + 1][3];
      }
      if (!last_stage) y[C - 1] = fmaf((st == 2) ? db[C - 1] : hb[C - 1], G_prev - GC, cn[C - 1]);
    }
    gh[C] = __shfl(gh[0], rsrc, 64);
    if (xr) gh[C] = xch[wave + 1][4];

    if constexpr (FOUT) {  // face mass flux per unit area: faces s0 .. s0+C-1, and n+1
      float *fo = a.fout + (size_t)(nx + 3) * ny * nz * b + vbase;
#pragma unroll
      for (int q = 0; q < C; ++q) {
        const int p = s0 + q;
        if (act && p <= n + 1) fo[p] = fmaf(a.fscale, gh[q], a.fout_add ? fo[p] : 0.0f);
      }
    }
    // Final update. Beyond cell n the factor fb is 0 and cn holds the east
    // ghost, so a lane whose cells end at n+2 or earlier stores them all
    // with wide stores (cells n+1, n+2 get the ghost value they hold); a
    // lane reaching further (the last segment) stores cells up to n only.
    // transfers through faces 1 and n+1: FP32 products for the compensated
    // state; for the FP32 state the raw face values, scaled in FP64 below
    float tl0 = 0.0f, trn = 0.0f;
    double fl0 = 1.0, fln = 1.0;
    float res = 0.0f;
    float hv[C], lv[C];
#pragma unroll
    for (int q = 0; q < C; ++q) {
      if constexpr (COMP) {
        const float tl = fb[q] * gh[q];
        const float tr = fb[q] * gh[q + 1];
        if (q == 0) tl0 = tl;
        if (q == qn) trn = tr;
        hv[q] = cn[q];
        lv[q] = lo[q];
        comp_add(hv[q], lv[q], tl);
        comp_add(hv[q], lv[q], -tr);
      } else {
        if (q == 0) {
          tl0 = gh[0];
          fl0 = f64[0];
        }
        if (q == qn) {
          trn = gh[q + 1];
          fln = f64[q];
        }
        hv[q] = final_update_fp64(cn[q], gh[q], gh[q + 1], f64[q], wq[q], true, res);
      }
    }
    if (act) {
      if (s0 + C - 1 <= n + 2) {
        store_cells<C>(a.chi + cb + s0, hv);
        if constexpr (COMP) store_cells<C>(a.clo + cb + s0, lv);
      } else {
#pragma unroll
        for (int q = 0; q < C; ++q) {
          if (s0 + q <= n) {
            a.chi[cb + s0 + q] = hv[q];
            if constexpr (COMP) a.clo[cb + s0 + q] = lv[q];
          }
        }
      }
    }
    if (act) {
      if (seg == 0) {
        acc_rb += (double)RFb * area * (double)dt;
        acc_lb -= (double)tl0 * fl0 * a.dxb64[1] * area;
      }
      if (qn >= -1 && qn <= C - 2) acc_re += (double)RFe * area * (double)dt;  // holds face n+1
      if (qn >= 0 && qn <= C - 1) {  // holds cell n
        acc_le += (double)trn * fln * a.dxb64[n] * area;
      }
      acc_res += (double)res * area;
    }
  }
  block_partials5(acc_rb, acc_re, acc_lb, acc_le, acc_res, a.partials, a.rpartials, (size_t)by * a.gx + bx);
}

template <int C, int LIM, bool COMP, int NST, bool FOUT = false>
__global__ void __launch_bounds__(256) k_sweep_x(SweepArgs a) {
  sweep_x_body<C, LIM, COMP, NST, FOUT>(a, blockIdx.x, blockIdx.y);
}

template <int C, int LIM, bool COMP, int NST>
__global__ void __launch_bounds__(256) k_sweep_x_batch(SweepBatch b) {
  const int m = batch_member(b, blockIdx.x);
  const SweepArgs a = member_args(b, m);
  const int lb = blockIdx.x - ((cint *)b.bstart)[m];
  sweep_x_body<C, LIM, COMP, NST>(a, lb % a.gx, lb / a.gx);
}

}  // namespace f3d
