// ads_sweep_impl.h -- the directional-split Kurganov-Tadmor sweep on gfx950.
//
// One kernel template serves the x, y and z sweeps of
// Sources/mod_ADS.F90:233-811. Each grid line of the sweep axis is split into
// S segments of C consecutive cells; one thread owns one segment and keeps
// its cells, the RK4 stage values and the RK-weighted face fluxes in VGPRs
// for the whole sweep, so the concentration is read once and written once
// per sweep. Between RK stages the two edge cells on each side of every
// segment are exchanged through LDS (double-buffered, one barrier per stage).
//
//   x sweep: consecutive lanes own consecutive segments of one x-line
//            (a 64-lane wave covers up to 64*C cells of a line).
//   y, z sweeps: consecutive lanes own the same segment of consecutive
//            x-columns, so every global load and store is coalesced along x;
//            the segments of one line are spread over the waves of the
//            workgroup.
//
// All bins of a line are processed by the same workgroup (bin loop inside
// the kernel), so the face winds and diffusion coefficients are read from
// HBM once per sweep, not once per bin.
//
// Numerics per face f (between cells f-1 and f), for stage input y:
//   d_f   = y_f - y_{f-1}
//   m_i   = phi(r_i) * d_{i+1}, r_i = d_i / d_{i+1}     (limiter slope, no
//           division for minmod/superbee; |d_{i+1}| < 1e-12 branch as in
//           Sources/mod_ADS.F90:1140-1146,1171-1176,1207-1212)
//   F_f   = max(u_f,0) (y_{f-1} + m_{f-1}/2) + min(u_f,0) (y_f - m_f/2)
//           (= KT flux with a = |u|, Sources/mod_ADS.F90:1103-1109)
//   P_f   = (k_{f-1} + k_f) / (2 dxp_{f-1}) * d_f           (:1111-1112)
//   G_f   = F_f - P_f,  RHS_i = (G_i - G_{i+1}) / dxb_i     (:1114)
// RK4 as in the reference (ghost cells frozen for the sweep, :330,:519,:727):
//   y1 = c + dt/2 RHS(c); y2 = c + dt/2 RHS(y1); y3 = c + dt RHS(y2)
//   c_new = c + dt/dxb_i (Gh_i - Gh_{i+1}),  Gh = (G1 + 2 G2 + 2 G3 + G4)/6
// (the reference's new_c accumulation rewritten in flux form; identical in
// exact arithmetic). Boundary fluxes for the reference accumulators use the
// last stage only (:394-396,:788-795); the conservation ledger uses Gh.
#pragma once

#include "ads_internal.h"

namespace f3d {

__device__ __forceinline__ size_t cidx(int i, int j, int k, int b, int nx, int ny, int nz) {
  return (size_t)(i + 1) + (size_t)(nx + 4) * ((size_t)(j + 1) + (size_t)(ny + 4) * ((size_t)(k + 1) + (size_t)(nz + 4) * (size_t)b));
}

// Limiter slope m = phi(d_i/d_{i+1}) * d_{i+1}.
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

// Compensated (hi, lo) += t. TwoSum of hi and t, error folded into lo, then
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

template <int DIR, int C, int LIM, bool COMP>
__global__ void __launch_bounds__(512) k_sweep(SweepArgs a) {
  static_assert(C >= 2, "segments need two edge cells");
  extern __shared__ float ex[];  // [2][4][blockDim] edge exchange
  __shared__ double red[4][8];

  const int t = threadIdx.x;
  const int nthr = blockDim.x;
  const int S = a.S;
  const int n = a.n;
  const int nx = a.nx, ny = a.ny, nz = a.nz;

  int seg, lslot, tA, tB;  // segment, line slot, transverse indices (1-based)
  bool line_ok;
  if constexpr (DIR == 0) {
    seg = t % S;
    lslot = t / S;
    const int lpb = nthr / S;
    const int line = blockIdx.x * lpb + lslot;
    line_ok = (lslot < lpb) && (line < ny * nz);
    tA = line % ny + 1;  // j
    tB = line / ny + 1;  // k
  } else {
    lslot = t % a.TX;
    seg = t / a.TX;
    const int tile = blockIdx.x % a.ntrans;
    const int other = blockIdx.x / a.ntrans;
    tA = tile * a.TX + lslot + 1;  // i
    tB = other + 1;                // k (y sweep) or j (z sweep)
    line_ok = (tA <= nx);
  }
  const bool act = line_ok && (seg < S);
  const int s0 = seg * C + 1;
  // Threads without a line (x-sweep tail, columns beyond nx) behave as a
  // one-segment line so they never read an exchange slot of another line.
  const bool first = (seg == 0) || !act;
  const bool last = (seg == S - 1) || !act;
  const int qn = n - s0;  // local index of cell n in the last segment

  // exchange slots: this thread writes slot t; neighbours along the line
  // are one slot (x sweep) or one row of TX slots (y, z sweeps) away
  const int left_ex = (DIR == 0) ? t - 1 : t - a.TX;
  const int right_ex = (DIR == 0) ? t + 1 : t + a.TX;

  // ---- geometry of this line ---------------------------------------------
  size_t cbase = 0, cstride = 1;  // element of position P (for bin 0) = cbase + P*cstride
  size_t cbin = (size_t)(nx + 4) * (ny + 4) * (nz + 4);
  size_t vbase = 0, vstride = 1;  // face f -> vbase + f*vstride
  size_t kbase = 0, kstride = 1;  // cell P -> kbase + (P+1)*kstride
  double area = 0.0;
  if (act) {
    if constexpr (DIR == 0) {
      cbase = cidx(0, tA, tB, 0, nx, ny, nz);
      cstride = 1;
      vbase = (size_t)(nx + 3) * ((size_t)(tA - 1) + (size_t)ny * (tB - 1));
      vstride = 1;
      kbase = (size_t)(nx + 4) * ((size_t)(tA - 1) + (size_t)ny * (tB - 1));
      kstride = 1;
      area = a.areaB[tB] * a.areaA[tA];  // dX3_b(k) * dX2_b(j)
    } else if constexpr (DIR == 1) {
      cbase = cidx(tA, 0, tB, 0, nx, ny, nz);
      cstride = (size_t)(nx + 4);
      vbase = (size_t)(tA - 1) + (size_t)nx * (size_t)(ny + 3) * (tB - 1);
      vstride = (size_t)nx;
      kbase = (size_t)(tA - 1) + (size_t)nx * (size_t)(ny + 4) * (tB - 1);
      kstride = (size_t)nx;
      area = a.areaB[tB] * a.areaA[tA];  // dX3_b(k) * dX1_b(i)
    } else {
      cbase = cidx(tA, tB, 0, 0, nx, ny, nz);
      cstride = (size_t)(nx + 4) * (ny + 4);
      vbase = (size_t)(tA - 1) + (size_t)nx * (tB - 1);
      vstride = (size_t)nx * ny;
      kbase = vbase;
      kstride = (size_t)nx * ny;
      area = a.areaB[tB] * a.areaA[tA];  // dX2_b(j) * dX1_b(i)
    }
  }

  // ---- bin-independent coefficients ---------------------------------------
  float vel[C + 1], kf[C + 1], ib[C];
  const float s = a.stime, s1 = 1.0f - a.stime;
#pragma unroll
  for (int q = 0; q <= C; ++q) {
    int f = s0 + q;
    f = f > n + 2 ? n + 2 : f;
    int km = f - 1 > n + 2 ? n + 2 : f - 1;
    float v = 0.0f, k0 = 0.0f, k1 = 0.0f, h = 0.0f;
    if (act) {
      v = s1 * a.vel1[vbase + (size_t)f * vstride] + s * a.vel2[vbase + (size_t)f * vstride];
      k0 = a.kd[kbase + (size_t)(km + 1) * kstride];
      k1 = a.kd[kbase + (size_t)(f + 1) * kstride];
      h = a.hdxp[f];
    }
    vel[q] = v;
    kf[q] = (k0 + k1) * h;
  }
#pragma unroll
  for (int q = 0; q < C; ++q) {
    int p = s0 + q;
    p = p > n + 2 ? n + 2 : p;
    ib[q] = act ? a.idxb[p] : 0.0f;
  }
  // boundary-face winds (face 1 for the first segment, face n+1 for the last)
  float vbeg = 0.0f, vend = 0.0f;
  if (act) {
    vbeg = s1 * a.vel1[vbase + vstride] + s * a.vel2[vbase + vstride];
    vend = s1 * a.vel1[vbase + (size_t)(n + 1) * vstride] + s * a.vel2[vbase + (size_t)(n + 1) * vstride];
  }
  const double dxb_first = a.dxb64[1], dxb_last = a.dxb64[n];

  float up[C + 1], um[C + 1];
  if constexpr (DIR != 2) {
#pragma unroll
    for (int q = 0; q <= C; ++q) {
      up[q] = fmaxf(vel[q], 0.0f);
      um[q] = fminf(vel[q], 0.0f);
    }
  }
  double acc_ref_beg = 0.0, acc_ref_end = 0.0, acc_led_beg = 0.0, acc_led_end = 0.0;
  int stage_ctr = 0;

  const int bfirst = a.bin0 + blockIdx.y * a.nbin;
  const int blast = min(bfirst + a.nbin, a.nbins);
  for (int b = bfirst; b < blast; ++b) {
    // velocities of this bin (w_eff = w - vs for the z sweep)
    float vb = vbeg, ve = vend;
    if constexpr (DIR == 2) {
      const size_t vsb = (size_t)(tA - 1) + (size_t)nx * ((size_t)(tB - 1) + (size_t)ny * (size_t)(nz + 3) * b);
      const size_t vss = (size_t)nx * ny;
#pragma unroll
      for (int q = 0; q <= C; ++q) {
        int f = s0 + q;
        f = f > n + 2 ? n + 2 : f;
        const float w = act ? vel[q] - a.vs[vsb + (size_t)f * vss] : 0.0f;
        up[q] = fmaxf(w, 0.0f);
        um[q] = fminf(w, 0.0f);
      }
      if (act) {
        vb = vbeg - a.vs[vsb + vss];
        ve = vend - a.vs[vsb + (size_t)(n + 1) * vss];
      }
    }

    // ---- load the segment ---------------------------------------------
    const size_t cb = cbase + cbin * (size_t)b;
    float cn[C], lo[C];
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const int p = s0 + q;
      float v = 0.0f, l = 0.0f;
      if (act && p <= n) {
        v = a.chi[cb + (size_t)p * cstride];
        if constexpr (COMP) l = a.clo[cb + (size_t)p * cstride];
      }
      cn[q] = v;
      lo[q] = l;
    }
    // free-flow ghosts (Sources/mod_ADS.F90:1239-1324), frozen for the sweep
    float gL = 0.0f, gR = 0.0f;
    if (first) gL = (vb > 0.0f) ? (2.0f * 0.0f - cn[0]) : cn[0];
    if (last) {
      float cl = cn[0];
#pragma unroll
      for (int q = 1; q < C; ++q) cl = (q == qn) ? cn[q] : cl;
      gR = (ve < 0.0f) ? (2.0f * 0.0f - cl) : cl;
#pragma unroll
      for (int q = 0; q < C; ++q) cn[q] = (q > qn) ? gR : cn[q];
    }
    if (act && a.write_ghosts) {
      if (first) {
        a.chi[cb - cstride] = gL;
        a.chi[cb] = gL;
        if constexpr (COMP) { a.clo[cb - cstride] = 0.0f; a.clo[cb] = 0.0f; }
      }
      if (last) {
        a.chi[cb + (size_t)(n + 1) * cstride] = gR;
        a.chi[cb + (size_t)(n + 2) * cstride] = gR;
        if constexpr (COMP) {
          a.clo[cb + (size_t)(n + 1) * cstride] = 0.0f;
          a.clo[cb + (size_t)(n + 2) * cstride] = 0.0f;
        }
      }
    }

    float y[C], gh[C + 1];
#pragma unroll
    for (int q = 0; q < C; ++q) y[q] = cn[q];
#pragma unroll
    for (int q = 0; q <= C; ++q) gh[q] = 0.0f;
    float Fb = 0.0f, Pb = 0.0f, Fe = 0.0f, Pe = 0.0f;

    const int nst = (a.tm == 1) ? 1 : 4;
    for (int st = 0; st < nst; ++st) {
      // edge exchange
      float *buf = ex + (size_t)(stage_ctr & 1) * 4 * nthr;
      ++stage_ctr;
      buf[0 * nthr + t] = y[0];
      buf[1 * nthr + t] = y[1];
      buf[2 * nthr + t] = y[C - 2];
      buf[3 * nthr + t] = y[C - 1];
      __syncthreads();
      float yl2 = gL, yl1 = gL, yr0 = gR, yr1 = gR;
      if (!first) {
        yl2 = buf[2 * nthr + left_ex];
        yl1 = buf[3 * nthr + left_ex];
      }
      if (!last) {
        yr0 = buf[0 * nthr + right_ex];
        yr1 = buf[1 * nthr + right_ex];
      }
      // window w[0..C+3] = positions s0-2 .. s0+C+1; faces are swept left to
      // right with rolling differences/slopes to keep register pressure low.
      float w[C + 4];
      w[0] = yl2;
      w[1] = yl1;
#pragma unroll
      for (int q = 0; q < C; ++q) w[q + 2] = y[q];
      w[C + 2] = yr0;
      w[C + 3] = yr1;
      const bool last_stage = (st == nst - 1);
      const float wt = (a.tm == 1) ? 1.0f : ((st == 0 || st == 3) ? (1.0f / 6.0f) : (1.0f / 3.0f));
      const bool upd = (a.tm != 1) && !last_stage;
      const float h = (st == 2) ? a.dt : 0.5f * a.dt;
      float d_cur = w[2] - w[1];                                 // face s0
      float m_prev = limiter_slope<LIM>(w[1] - w[0], d_cur);     // cell s0-1
      float G_prev = 0.0f;
#pragma unroll
      for (int q = 0; q <= C; ++q) {
        const float d_next = w[q + 3] - w[q + 2];                // face s0+q+1
        const float m_cur = limiter_slope<LIM>(d_cur, d_next);   // cell s0+q
        const float cL = fmaf(0.5f, m_prev, w[q + 1]);
        const float cR = fmaf(-0.5f, m_cur, w[q + 2]);
        const float F = fmaf(up[q], cL, um[q] * cR);
        const float P = kf[q] * d_cur;
        const float G = F - P;
        if (last_stage) {
          if (q == 0) { Fb = F; Pb = P; }
          if (q == qn + 1) { Fe = F; Pe = P; }
        }
        gh[q] = fmaf(wt, G, gh[q]);
        if (q > 0 && upd) {  // next stage input of cell s0+q-1
          const float yn = fmaf(h, (G_prev - G) * ib[q - 1], cn[q - 1]);
          y[q - 1] = (last && q - 1 > qn) ? gR : yn;
        }
        G_prev = G;
        d_cur = d_next;
        m_prev = m_cur;
      }
    }

    // ---- final update, store ----------------------------------------------
    const float dt = a.dt;
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
          float hi = cn[q], l = lo[q];
          comp_add(hi, l, tl);
          comp_add(hi, l, -tr);
          a.chi[cb + (size_t)p * cstride] = hi;
          a.clo[cb + (size_t)p * cstride] = l;
        } else {
          a.chi[cb + (size_t)p * cstride] = fmaf(dt * (gh[q] - gh[q + 1]), ib[q], cn[q]);
        }
      }
    }
    if (act) {
      if (first) {
        const float fr = Fb + Pb;
        acc_ref_beg += (double)fr * area * (double)dt;
        acc_led_beg -= (double)tl0 * dxb_first * area;
        if constexpr (DIR == 2) {
          const size_t ai = (size_t)(tA + 1) + (size_t)(nx + 4) * ((size_t)(tB + 1) + (size_t)(ny + 4) * b);
          a.acum[ai] = a.acum[ai] + fabs((double)fr) * (double)dt;
        }
      }
      if (last) {
        const float fr = Fe + Pe;
        acc_ref_end += (double)fr * area * (double)dt;
        acc_led_end += (double)trn * dxb_last * area;
      }
    }
  }

  // ---- deterministic block reduction of the boundary sums ------------------
  double v4[4] = {acc_ref_beg, acc_ref_end, acc_led_beg, acc_led_end};
  const int lane = t & 63, wid = t >> 6;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const double r = wave_sum(v4[e]);
    if (lane == 0) red[e][wid] = r;
  }
  __syncthreads();
  if (t < 4) {
    double sum = 0.0;
    const int nw = (nthr + 63) >> 6;
    for (int wv = 0; wv < nw; ++wv) sum += red[t][wv];
    a.partials[((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 4 + t] = sum;
  }
}

}  // namespace f3d
