// ads_sweep_impl.h -- the directional-split Kurganov-Tadmor sweeps on gfx950.
//
// Sources/mod_ADS.F90:233-811 integrate, for every grid line of the sweep
// axis and every bin, one RK4 (or Euler) step of the 1-D KT scheme with the
// free-flow ghosts frozen for the whole sweep. Two kernels do that here:
//
//  k_stream  (y and z sweeps): one thread per line; lanes of a wave hold 64
//    consecutive x-columns, so every load/store is a coalesced 256-byte row.
//    The thread walks its line once; the four RK stages form a software
//    pipeline (stage s lags stage s-1 by two cells, the width of the KT
//    stencil), so each cell is read once, written once, and no LDS or
//    barrier is needed. Small register rings hold the last cells, face winds
//    and face fluxes; inputs are prefetched kPF cells ahead. The waves of a
//    workgroup are bins of the same columns, so the wind/diffusion rows they
//    all read are shared through L1.
//
//  k_sweep_x (x sweep): lines are contiguous in memory, so one wave owns a
//    line (or several short ones); each lane keeps C consecutive cells, the
//    stage values and the face fluxes in VGPRs and trades two edge cells
//    with its neighbours per stage through cross-lane permutes (no LDS, no
//    barrier). The next bin's cells are prefetched while the current bin is
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

#include "ads_internal.h"

namespace f3d {

__device__ __forceinline__ size_t cidx(int i, int j, int k, int b, int nx, int ny, int nz) {
  return (size_t)(i + 1) + (size_t)(nx + 4) * ((size_t)(j + 1) + (size_t)(ny + 4) * ((size_t)(k + 1) + (size_t)(nz + 4) * (size_t)b));
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

// =============================================================================
// k_stream: y (DIR 1) and z (DIR 2) sweeps, one thread per line
// =============================================================================

// Raw buffer access with hardware range checking: a load outside
// [0, bytes) returns 0 and a store there is dropped, so the pipeline body
// needs no branches for edge cells, idle lanes or overshooting iterations.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
constexpr uint32_t kOOB = 0x80000000u;  // beyond every buffer used here (< 2 GiB each)

__device__ __forceinline__ rsrc_t make_rsrc(const void *base, size_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float bload(rsrc_t r, uint32_t byte_off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)byte_off, 0, 0));
}
__device__ __forceinline__ void bstore(rsrc_t r, uint32_t byte_off, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)byte_off, 0, 0);
}

constexpr int kPF = 4;  // prefetch distance (cells)

template <int DIR, int LIM, bool COMP, int NST>
__global__ void __launch_bounds__(512) k_stream(SweepArgs a) {
  // the wave index is uniform across the wave: say so, so that the per-wave
  // buffer descriptors live in SGPRs
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nx = a.nx, ny = a.ny, nz = a.nz, n = a.n;
  // block -> (bin group, x tile, transverse line index); wave -> bin
  const int groups = a.ntrans2;
  const int g = blockIdx.x % groups;
  const int rest = blockIdx.x / groups;
  const int tile = rest % a.ntrans;
  const int t2 = rest / a.ntrans + 1;  // k (y sweep) or j (z sweep)
  const int i = tile * 64 + lane + 1;
  const int b = g * a.nbin + wave;
  const bool wave_ok = (wave < a.nbin) && (b < a.nbins);
  const bool act = wave_ok && (i <= nx);
  const int bb = wave_ok ? b : 0;

  const size_t cbin = (size_t)(nx + 4) * (ny + 4) * (nz + 4);
  const rsrc_t rc = make_rsrc(a.chi + cbin * bb, cbin * 4);
  const rsrc_t rl = make_rsrc(COMP ? a.clo + cbin * bb : a.chi + cbin * bb, cbin * 4);
  const size_t nvel = (DIR == 1) ? (size_t)nx * (ny + 3) * nz : (size_t)nx * ny * (nz + 3);
  const size_t nkd = (DIR == 1) ? (size_t)nx * (ny + 4) * nz : (size_t)nx * ny * (nz + 4);
  const rsrc_t rv1 = make_rsrc(a.vel1, nvel * 4);
  const rsrc_t rv2 = make_rsrc(a.vel2, nvel * 4);
  const rsrc_t rk = make_rsrc(a.kd, nkd * 4);
  const size_t nvs = (size_t)nx * ny * (nz + 3);
  const rsrc_t rvs = make_rsrc(DIR == 2 ? a.vs + nvs * bb : a.vel1, (DIR == 2 ? nvs : nvel) * 4);
  const rsrc_t rh = make_rsrc(a.hdxp, (size_t)(n + 3) * 4);
  const rsrc_t rib = make_rsrc(a.idxb, (size_t)(n + 3) * 4);

  // element offsets of position/face 0 of this lane's line, and strides
  uint32_t cb0, cs, vb, vst, kb, kst;
  if constexpr (DIR == 1) {
    cb0 = (uint32_t)((i + 1) + (nx + 4) * (1 + (size_t)(ny + 4) * (t2 + 1)));
    cs = (uint32_t)(nx + 4);
    vb = (uint32_t)((i - 1) + (size_t)nx * (ny + 3) * (t2 - 1));
    vst = (uint32_t)nx;
    kb = (uint32_t)((i - 1) + (size_t)nx * (ny + 4) * (t2 - 1));
    kst = (uint32_t)nx;
  } else {
    cb0 = (uint32_t)((i + 1) + (nx + 4) * ((size_t)(t2 + 1) + (size_t)(ny + 4)));
    cs = (uint32_t)((nx + 4) * (ny + 4));
    vb = (uint32_t)((i - 1) + (size_t)nx * (t2 - 1));
    vst = (uint32_t)(nx * ny);
    kb = vb;
    kst = vst;
  }
  const double area = act ? a.areaB[t2] * a.areaA[i] : 0.0;  // dX3_b(k) dX1_b(i) / dX2_b(j) dX1_b(i)
  const float s = a.stime, s1 = 1.0f - a.stime;
  const float dt = a.dt;

  auto face_v = [&](int f) -> float {  // wind (w_eff for z) at face f of this line
    float v = s1 * bload(rv1, (vb + (uint32_t)f * vst) * 4u) + s * bload(rv2, (vb + (uint32_t)f * vst) * 4u);
    if constexpr (DIR == 2) v -= bload(rvs, (vb + (uint32_t)f * vst) * 4u);
    return v;
  };

  // free-flow ghosts from cells 1, n and faces 1, n+1 (Sources/mod_ADS.F90:1239-1324)
  const float c1 = bload(rc, (cb0 + cs) * 4u);
  const float cnn = bload(rc, (cb0 + (uint32_t)n * cs) * 4u);
  const float vfirst = face_v(1), vlast = face_v(n + 1);
  const float gL = (vfirst > 0.0f) ? (2.0f * 0.0f - c1) : c1;
  const float gR = (vlast < 0.0f) ? (2.0f * 0.0f - cnn) : cnn;
  {
    const bool wg = act && a.write_ghosts;
    bstore(rc, wg ? (cb0 - cs) * 4u : kOOB, gL);
    bstore(rc, wg ? cb0 * 4u : kOOB, gL);
    bstore(rc, wg ? (cb0 + (uint32_t)(n + 1) * cs) * 4u : kOOB, gR);
    bstore(rc, wg ? (cb0 + (uint32_t)(n + 2) * cs) * 4u : kOOB, gR);
    if constexpr (COMP) {
      bstore(rl, wg ? (cb0 - cs) * 4u : kOOB, 0.0f);
      bstore(rl, wg ? cb0 * 4u : kOOB, 0.0f);
      bstore(rl, wg ? (cb0 + (uint32_t)(n + 1) * cs) * 4u : kOOB, 0.0f);
      bstore(rl, wg ? (cb0 + (uint32_t)(n + 2) * cs) * 4u : kOOB, 0.0f);
    }
  }

  // rings: cells (16), low words (8, fetched kPF ahead of the final update),
  // raw face data (8) and face coefficients (8)
  float cn[16], lo[8];
  float rw1[8], rw2[8], rws[8], rkk[8];
  float vf[8], kf[8], gh[8];
#pragma unroll
  for (int q = 0; q < 16; ++q) cn[q] = 0.0f;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    rw1[q] = rw2[q] = rws[q] = rkk[q] = 0.0f;
    vf[q] = kf[q] = gh[q] = lo[q] = 0.0f;
  }
  // per-stage pipeline state: previous two inputs, last difference, last
  // slope, last face flux
  float sa1[NST], sa2[NST], sd[NST], sm[NST], sG[NST];
#pragma unroll
  for (int st = 0; st < NST; ++st) sa1[st] = sa2[st] = sd[st] = sm[st] = sG[st] = gL;

  float Fb = 0.0f, Pb = 0.0f, Fe = 0.0f, Pe = 0.0f, tl1 = 0.0f, trn = 0.0f;
  const int tend = n + 2 * NST + 1;  // iteration of the last final output (cell n)

  for (int t0 = -16; t0 <= tend; t0 += 16) {
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int t = t0 + u;
      // ---- prefetch cell t+kPF-1 and face t+kPF-2 (edge/overshoot reads are
      // harmless: in-buffer cells of other lines or 0 from the range check)
      {
        const int P = t + kPF - 1;
        cn[(u + kPF - 1) & 15] = bload(rc, (cb0 + (uint32_t)P * cs) * 4u);
        if constexpr (COMP) {
          const int Q = t + kPF - 2 * NST - 1;  // cell finalised kPF iterations from now
          lo[(u + kPF - 2 * NST - 1) & 7] = bload(rl, (cb0 + (uint32_t)Q * cs) * 4u);
        }
        const int f = t + kPF - 2;
        const uint32_t fo = (vb + (uint32_t)f * vst) * 4u;
        rw1[(u + kPF - 2) & 7] = bload(rv1, fo);
        rw2[(u + kPF - 2) & 7] = bload(rv2, fo);
        if constexpr (DIR == 2) rws[(u + kPF - 2) & 7] = bload(rvs, fo);
        rkk[(u + kPF - 2) & 7] = bload(rk, (kb + (uint32_t)(f + 1) * kst) * 4u);
      }
      // ---- coefficients of face t-2 (first used by stage 1 in this iteration)
      {
        const int f = t - 2;
        const int sl = (u - 2) & 7;
        float v = s1 * rw1[sl] + s * rw2[sl];
        if constexpr (DIR == 2) v -= rws[sl];
        vf[sl] = v;
        kf[sl] = (rkk[(u - 3) & 7] + rkk[sl]) * bload(rh, (uint32_t)f * 4u);
      }
      // ---- the pipeline: stage st consumes position t+1-2(st+1)
      float v;
      {
        const int P = t - 1;
        v = (P <= 0) ? gL : ((P > n) ? gR : cn[(u - 1) & 15]);
      }
#pragma unroll
      for (int st = 0; st < NST; ++st) {
        const int f = t - 2 * (st + 1);  // face completed by this stage
        const int fs = (u - 2 * (st + 1)) & 7;
        const float d = v - sa1[st];
        const float m = limiter_slope<LIM>(sd[st], d);
        const float cL = fmaf(0.5f, sm[st], sa2[st]);
        const float cR = fmaf(-0.5f, m, sa1[st]);
        const float uf = vf[fs];
        const float F = uf * ((uf > 0.0f) ? cL : cR);  // upwind = KT flux with a = |u|
        const float Pd = kf[fs] * sd[st];
        const float G = F - Pd;
        float wgt;
        if constexpr (NST == 1)
          wgt = 1.0f;
        else
          wgt = (st == 0 || st == 3) ? (1.0f / 6.0f) : (1.0f / 3.0f);
        gh[fs] = (st == 0) ? wgt * G : fmaf(wgt, G, gh[fs]);
        if (st == NST - 1) {
          Fb = (f == 1) ? F : Fb;
          Pb = (f == 1) ? Pd : Pb;
          Fe = (f == n + 1) ? F : Fe;
          Pe = (f == n + 1) ? Pd : Pe;
        }
        const int q = f - 1;  // output cell of this stage
        const float Gp = sG[st];
        sa2[st] = sa1[st];
        sa1[st] = v;
        sd[st] = d;
        sm[st] = m;
        sG[st] = G;
        if (st < NST - 1) {
          const float h = (st == 2) ? dt : 0.5f * dt;
          const float yn = fmaf(h * bload(rib, (uint32_t)q * 4u), Gp - G, cn[(u - 2 * (st + 1) - 1) & 15]);
          v = (q <= 0) ? gL : ((q > n) ? gR : yn);  // next stage input (position q)
        }
      }
      // ---- final update of cell q = t - 2 NST - 1
      {
        const int q = t - 2 * NST - 1;
        const bool qok = act && q >= 1 && q <= n;
        const uint32_t off = qok ? (cb0 + (uint32_t)q * cs) * 4u : kOOB;
        const int cl = (u - 2 * NST - 1) & 15;
        const float ib = bload(rib, (uint32_t)q * 4u);
        const float gq = gh[(u - 2 * NST - 1) & 7], gq1 = gh[(u - 2 * NST) & 7];
        const float tl = (dt * gq) * ib;
        const float tr = (dt * gq1) * ib;
        tl1 = (q == 1) ? tl : tl1;
        trn = (q == n) ? tr : trn;
        if constexpr (COMP) {
          float hv = cn[cl], lv = lo[(u - 2 * NST - 1) & 7];
          comp_add(hv, lv, tl);
          comp_add(hv, lv, -tr);
          bstore(rc, off, hv);
          bstore(rl, off, lv);
        } else {
          bstore(rc, off, fmaf(dt * (gq - gq1), ib, cn[cl]));
        }
      }
    }
  }

  double rb = 0.0, re = 0.0, lb = 0.0, le = 0.0;
  if (act) {
    const float frb = Fb + Pb, fre = Fe + Pe;
    rb = (double)frb * area * (double)dt;
    re = (double)fre * area * (double)dt;
    lb = -(double)tl1 * a.dxb64[1] * area;
    le = (double)trn * a.dxb64[n] * area;
    if constexpr (DIR == 2) {
      const size_t ai = (size_t)(i + 1) + (size_t)(nx + 4) * ((size_t)(t2 + 1) + (size_t)(ny + 4) * bb);
      a.acum[ai] = a.acum[ai] + fabs((double)frb) * (double)dt;
    }
  }
  block_partials(rb, re, lb, le, a.partials, blockIdx.x);
}

// =============================================================================
// k_sweep_x: x sweep, one wave per line, C cells per lane
// =============================================================================

template <int C, int LIM, bool COMP, int NST>
__global__ void __launch_bounds__(256) k_sweep_x(SweepArgs a) {
  static_assert(C >= 2, "segments need two edge cells");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int S = a.S, n = a.n;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int lpw = 64 / S;  // lines per wave
  const int lslot = lane / S, seg = lane % S;
  const int line = (blockIdx.x * (blockDim.x >> 6) + wave) * lpw + lslot;
  const bool act = (lslot < lpw) && (line < ny * nz);
  const int j = act ? line % ny + 1 : 1;
  const int k = act ? line / ny + 1 : 1;
  const int s0 = seg * C + 1;
  const bool first = (seg == 0) || !act;
  const bool last = (seg == S - 1) || !act;
  const int qn = n - s0;  // local index of cell n in the last segment
  const int lsrc = (lane + 63) & 63, rsrc = (lane + 1) & 63;

  const size_t cbase = cidx(0, j, k, 0, nx, ny, nz);
  const size_t cbin = (size_t)(nx + 4) * (ny + 4) * (nz + 4);
  const size_t vbase = (size_t)(nx + 3) * ((size_t)(j - 1) + (size_t)ny * (k - 1));
  const size_t kbase = (size_t)(nx + 4) * ((size_t)(j - 1) + (size_t)ny * (k - 1));
  const double area = a.areaB[k] * a.areaA[j];  // dX3_b(k) dX2_b(j)
  const float s = a.stime, s1 = 1.0f - a.stime;
  const float dt = a.dt;

  // wind and diffusion of this line, read once for all bins
  float up[C + 1], um[C + 1], kf[C + 1], ib[C];
  float vbeg = 0.0f, vend = 0.0f;
  {
    float kprev = act ? a.kd[kbase + (size_t)s0] : 0.0f;  // cell s0-1 (stored at index s0)
#pragma unroll
    for (int q = 0; q <= C; ++q) {
      int f = s0 + q;
      f = f > n + 2 ? n + 2 : f;
      float v = 0.0f, kc = 0.0f, h = 0.0f;
      if (act) {
        v = s1 * a.vel1[vbase + f] + s * a.vel2[vbase + f];
        kc = a.kd[kbase + (size_t)(f + 1)];
        h = a.hdxp[f];
      }
      up[q] = fmaxf(v, 0.0f);
      um[q] = fminf(v, 0.0f);
      kf[q] = (kprev + kc) * h;
      kprev = kc;
    }
#pragma unroll
    for (int q = 0; q < C; ++q) {
      int p = s0 + q;
      p = p > n + 2 ? n + 2 : p;
      ib[q] = act ? a.idxb[p] : 0.0f;
    }
    if (act) {
      vbeg = s1 * a.vel1[vbase + 1] + s * a.vel2[vbase + 1];
      vend = s1 * a.vel1[vbase + (size_t)(n + 1)] + s * a.vel2[vbase + (size_t)(n + 1)];
    }
  }

  double acc_rb = 0.0, acc_re = 0.0, acc_lb = 0.0, acc_le = 0.0;
  const int bfirst = a.bin0 + blockIdx.y * a.nbin;
  const int blast = min(bfirst + a.nbin, a.nbins);

  // prefetch the first bin
  float nxt[C], nlo[C];
#pragma unroll
  for (int q = 0; q < C; ++q) {
    const int p = s0 + q;
    nxt[q] = 0.0f;
    nlo[q] = 0.0f;
    if (act && p <= n && bfirst < blast) {
      nxt[q] = a.chi[cbase + cbin * bfirst + p];
      if constexpr (COMP) nlo[q] = a.clo[cbase + cbin * bfirst + p];
    }
  }

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
#pragma unroll
      for (int q = 0; q < C; ++q) {
        const int p = s0 + q;
        if (act && p <= n) {
          nxt[q] = a.chi[cb + cbin + p];
          if constexpr (COMP) nlo[q] = a.clo[cb + cbin + p];
        }
      }
    }
    // free-flow ghosts (Sources/mod_ADS.F90:1239-1265), frozen for the sweep
    float gL = 0.0f, gR = 0.0f;
    if (first) gL = (vbeg > 0.0f) ? (2.0f * 0.0f - cn[0]) : cn[0];
    if (last) {
      float cl = cn[0];
#pragma unroll
      for (int q = 1; q < C; ++q) cl = (q == qn) ? cn[q] : cl;
      gR = (vend < 0.0f) ? (2.0f * 0.0f - cl) : cl;
#pragma unroll
      for (int q = 0; q < C; ++q) cn[q] = (q > qn) ? gR : cn[q];
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
    float Fb = 0.0f, Pb = 0.0f, Fe = 0.0f, Pe = 0.0f;

#pragma unroll
    for (int st = 0; st < NST; ++st) {
      // neighbours' edge cells (same wave: cross-lane permute, no barrier)
      float yl2 = __shfl(y[C - 2], lsrc, 64);
      float yl1 = __shfl(y[C - 1], lsrc, 64);
      float yr0 = __shfl(y[0], rsrc, 64);
      float yr1 = __shfl(y[1], rsrc, 64);
      if (first) {
        yl2 = gL;
        yl1 = gL;
      }
      if (last) {
        yr0 = gR;
        yr1 = gR;
      }
      float w[C + 4];
      w[0] = yl2;
      w[1] = yl1;
#pragma unroll
      for (int q = 0; q < C; ++q) w[q + 2] = y[q];
      w[C + 2] = yr0;
      w[C + 3] = yr1;
      const bool last_stage = (st == NST - 1);
      float wgt;
      if constexpr (NST == 1)
        wgt = 1.0f;
      else
        wgt = (st == 0 || st == 3) ? (1.0f / 6.0f) : (1.0f / 3.0f);
      const float h = (st == 2) ? dt : 0.5f * dt;
      float d_cur = w[2] - w[1];
      float m_prev = limiter_slope<LIM>(w[1] - w[0], d_cur);
      float G_prev = 0.0f;
#pragma unroll
      for (int q = 0; q <= C; ++q) {
        const float d_next = w[q + 3] - w[q + 2];
        const float m_cur = limiter_slope<LIM>(d_cur, d_next);
        const float cL = fmaf(0.5f, m_prev, w[q + 1]);
        const float cR = fmaf(-0.5f, m_cur, w[q + 2]);
        const float F = fmaf(up[q], cL, um[q] * cR);
        const float P = kf[q] * d_cur;
        const float G = F - P;
        if (last_stage) {
          if (q == 0) {
            Fb = F;
            Pb = P;
          }
          if (q == qn + 1) {
            Fe = F;
            Pe = P;
          }
        }
        gh[q] = (st == 0) ? wgt * G : fmaf(wgt, G, gh[q]);
        if (q > 0 && !last_stage) {
          const float yn = fmaf(h, (G_prev - G) * ib[q - 1], cn[q - 1]);
          y[q - 1] = (last && q - 1 > qn) ? gR : yn;
        }
        G_prev = G;
        d_cur = d_next;
        m_prev = m_cur;
      }
    }

    float tl0 = 0.0f, trn = 0.0f;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const int p = s0 + q;
      const float tl = (dt * gh[q]) * ib[q];
      const float tr = (dt * gh[q + 1]) * ib[q];
      if (q == 0) tl0 = tl;
      if (q == qn) trn = tr;
      if (act && p <= n) {
        if constexpr (COMP) {
          float hv = cn[q], lv = lo[q];
          comp_add(hv, lv, tl);
          comp_add(hv, lv, -tr);
          a.chi[cb + p] = hv;
          a.clo[cb + p] = lv;
        } else {
          a.chi[cb + p] = fmaf(dt * (gh[q] - gh[q + 1]), ib[q], cn[q]);
        }
      }
    }
    if (act) {
      if (seg == 0) {
        acc_rb += (double)(Fb + Pb) * area * (double)dt;
        acc_lb -= (double)tl0 * a.dxb64[1] * area;
      }
      if (seg == S - 1) {
        acc_re += (double)(Fe + Pe) * area * (double)dt;
        acc_le += (double)trn * a.dxb64[n] * area;
      }
    }
  }
  block_partials(acc_rb, acc_re, acc_lb, acc_le, a.partials, (size_t)blockIdx.y * gridDim.x + blockIdx.x);
}

}  // namespace f3d
