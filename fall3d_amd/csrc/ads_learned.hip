// ads_learned.hip -- learned flux correction of the coarse-grid KT step.
//
// No counterpart in the FALL3D reference. A per-face network (ads_mlp.hip)
// corrects the horizontal face mass fluxes of a 2x horizontally coarsened
// grid. Plain KT stays the method and the fallback: with no network set, a
// context runs the plain step; with one, the step is
//   1. corrections from the state at the start of the step,
//   2. the plain step (source, x/y/z sweeps) unchanged,
//   3. c -= (dPhi(f+1) - dPhi(f)) / dxb on the x and y faces (conservative).
//
// Face features (32 bf16; also built for training rows): the 3x3x3 stencil
// of concentrations around the face's upwind cell, mirrored so that the
// normal flow and the tangential flow are non-negative, divided by its
// largest magnitude S; then the Courant numbers |u_n| dt/dxp, |u_t| dt/dx_t,
// w dt/dz and v_s dt/dz; one zero pad. The network output y gives
//   dPhi = sgn_n * S * dxp * 0.01 y   (mass per unit face area over the step),
// so the correction scales with the field (the transport is linear in c).
//
// Only faces in active tiles are corrected: a tile is 16 consecutive faces
// along x, active when one of its faces has a cell above the threshold. One
// step runs: live-cell bits (one read of c), tile flags and a list of the
// active tiles per axis, the network on the listed tiles, a list of the row
// segments next to an active tile, and the update of those segments' cells.
//
// Training pairs come from the solver itself: a fine run and a coarse run
// from the coarse-grained fine state, both with the sweeps' face-flux output
// on; the target of a coarse face is (fine flux averaged over its two fine
// faces and two fine steps - coarse flux) / (sgn_n S dxp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "ads_context.h"
#include "ads_mlp.h"
#include "ads_mlp_device.h"

using namespace f3d;
using namespace f3d::host;

struct fall3d_mlp {
  std::string err = "no error";
  hipStream_t stream = nullptr;
  float *params = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *part = nullptr;
  int part_waves = 0;
  PackedWeights *packed = nullptr;
  int adam_t = 0;
  // training rows on the device
  uint16_t *x = nullptr;
  float *t = nullptr;
  long cap = 0;
  unsigned int *dcount = nullptr;  // device row counter (collection)
  long count = 0;
  int4 *ref = nullptr;  // per row: face (axis | bin << 1, f, transverse, k)
  float *den = nullptr; // per row: sgn S dxp
  uint32_t *bidx = nullptr;
  long bidx_cap = 0;
  float *loss = nullptr;
  int loss_cap = 0;
};

namespace f3d {

// ---------------------------------------------------------------------------
// face geometry and features
// ---------------------------------------------------------------------------
struct FaceArgs {
  const float *c;  // state, bin 0 (reference layout with ghosts)
  int nx, ny, nz, nb;
  int axis;  // 0: x faces, 1: y faces
  const float *u1, *u2, *v1, *v2, *w1, *w2, *vs;
  float s, s1, dt;
  const float *hdxp[3];  // 0.5/dxp(f-1) at face f
  const float *idxb[3];  // 1/dxb(p) at cell p
  float smin;            // faces with S <= smin are inactive
  const float *smin_dev; // or, when set, the threshold in device memory
  float gain;            // fraction of the predicted correction applied
};

__device__ __forceinline__ size_t cix(const FaceArgs &a, int i, int j, int k, int b) {
  return (size_t)(i + 1) + (size_t)state_pitch(a.nx) * ((size_t)(j + 1) + (size_t)(a.ny + 4) * ((size_t)(k + 1) + (size_t)(a.nz + 4) * b));
}
__device__ __forceinline__ int clampi(int x, int lo, int hi) { return x < lo ? lo : (x > hi ? hi : x); }
__device__ __forceinline__ float lerpv(const FaceArgs &a, const float *p1, const float *p2, size_t i) {
  return a.s1 * p1[i] + a.s * p2[i];
}

// Interior face (axis, f in 2..n_axis, cross index tr, level k, bin b): the
// cell (i, j) of the face's L side is (f-1, tr) on x, (tr, f-1) on y.
struct FaceGeo {
  int iL, jL, iR, jR;  // L and R cells
  float sgn, tsgn;     // mirror signs (normal, tangential)
  int iU, jU;          // upwind cell
  float cn, ct, cw, cs;
  float dxp;
};

__device__ __forceinline__ FaceGeo face_geo(const FaceArgs &a, int f, int tr, int k, int b) {
  FaceGeo g;
  const int nx = a.nx, ny = a.ny;
  float un, ut;
  if (a.axis == 0) {
    g.iL = f - 1, g.jL = tr, g.iR = f, g.jR = tr;
    un = lerpv(a, a.u1, a.u2, (size_t)f + (size_t)(nx + 3) * ((size_t)(tr - 1) + (size_t)ny * (k - 1)));
    ut = 0.0f;
#pragma unroll
    for (int di = 0; di < 2; ++di)
#pragma unroll
      for (int dj = 0; dj < 2; ++dj)
        ut += lerpv(a, a.v1, a.v2, (size_t)(f - 1 + di - 1) + (size_t)nx * ((size_t)(tr + dj) + (size_t)(ny + 3) * (k - 1)));
  } else {
    g.iL = tr, g.jL = f - 1, g.iR = tr, g.jR = f;
    un = lerpv(a, a.v1, a.v2, (size_t)(tr - 1) + (size_t)nx * ((size_t)f + (size_t)(ny + 3) * (k - 1)));
    ut = 0.0f;
#pragma unroll
    for (int dj = 0; dj < 2; ++dj)
#pragma unroll
      for (int di = 0; di < 2; ++di)
        ut += lerpv(a, a.u1, a.u2, (size_t)(tr + di) + (size_t)(nx + 3) * ((size_t)(f - 1 + dj - 1) + (size_t)ny * (k - 1)));
  }
  ut *= 0.25f;
  float wv = 0.0f;
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const int i = side ? g.iR : g.iL, j = side ? g.jR : g.jL;
#pragma unroll
    for (int dk = 0; dk < 2; ++dk) wv += lerpv(a, a.w1, a.w2, (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (k + dk)));
  }
  wv *= 0.25f;
  g.sgn = (un >= 0.0f) ? 1.0f : -1.0f;
  g.tsgn = (ut >= 0.0f) ? 1.0f : -1.0f;
  g.iU = (un >= 0.0f) ? g.iL : g.iR;
  g.jU = (un >= 0.0f) ? g.jL : g.jR;
  const int ta = 1 - a.axis;  // transverse axis
  const int tp = (a.axis == 0) ? g.jU : g.iU;
  g.dxp = 0.5f / a.hdxp[a.axis][f];
  g.cn = fabsf(un) * a.dt / g.dxp;
  g.ct = fabsf(ut) * a.dt * a.idxb[ta][tp];
  g.cw = wv * a.dt * a.idxb[2][k];
  g.cs = a.vs[(size_t)(g.iU - 1) + (size_t)nx * ((size_t)(g.jU - 1) + (size_t)ny * ((size_t)k + (size_t)(a.nz + 3) * b))] *
         a.dt * a.idxb[2][k];
  return g;
}

// stencil value e (0..26) of the face: offsets (da, dt, dk) in -1..1 from the
// upwind cell, da along the (mirrored) normal, dt along the (mirrored)
// transverse axis; clamped to the interior
__device__ __forceinline__ float stencil(const FaceArgs &a, const FaceGeo &g, int e, int k, int b) {
  const int da = e % 3 - 1, dt = (e / 3) % 3 - 1, dk = e / 9 - 1;
  const int sa = (int)g.sgn * da, st = (int)g.tsgn * dt;
  int i, j;
  if (a.axis == 0) {
    i = g.iU + sa;
    j = g.jU + st;
  } else {
    i = g.iU + st;
    j = g.jU + sa;
  }
  i = clampi(i, 1, a.nx);
  j = clampi(j, 1, a.ny);
  const int kk = clampi(k + dk, 1, a.nz);
  return a.c[cix(a, i, j, kk, b)];
}

constexpr float kSmallest = 1e-30f;
// the network predicts kOutScale^-1 times the normalised correction (the
// corrections are ~1e-2 of S dxp; O(1) outputs train better)
constexpr float kOutScale = 0.01f;

// Lane (group g = l/16, column l%16) builds features 8g..8g+7 of its face as
// the B operand of the first layer; returns S (same in all groups).
__device__ __forceinline__ bf16x8 face_features(const FaceArgs &a, const FaceGeo &geo, bool ok, int k, int b,
                                                float &S) {
  const int g = (threadIdx.x & 63) >> 4;
  float val[8];
  float mx = 0.0f;
#pragma unroll
  for (int s = 0; s < 8; ++s) {
    const int e = 8 * g + s;
    float v = 0.0f;
    if (ok) {
      if (e < 27) {
        v = stencil(a, geo, e, k, b);
        mx = fmaxf(mx, fabsf(v));
      } else if (e == 27) {
        v = geo.cn;
      } else if (e == 28) {
        v = geo.ct;
      } else if (e == 29) {
        v = geo.cw;
      } else if (e == 30) {
        v = geo.cs;
      }
    }
    val[s] = v;
  }
  mx = fmaxf(mx, __shfl_xor(mx, 16, 64));
  mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
  // stencils below kSmallest (denormal range: 1/S would overflow) are
  // treated as empty: features zero, face inactive
  S = (mx > kSmallest) ? mx : 0.0f;
  bf16x8 xb;
#pragma unroll
  for (int s = 0; s < 8; ++s) xb[s] = (__bf16)((8 * g + s < 27) ? (S > 0.0f ? val[s] / S : 0.0f) : val[s]);
  return xb;
}

// tiles of 16 faces along x per row (a multiple of 4: rows start on whole
// 64-face groups, the tiles past the last face are empty) and rows
__device__ __host__ __forceinline__ void face_space(int axis, int nx, int ny, int nz, int nb, int &nt, long &rows) {
  if (axis == 0) {  // faces f = 2..nx along x; rows (j, k, b)
    nt = (nx - 1 + 63) / 64 * 4;
    rows = (long)ny * nz * nb;
  } else {  // cells i = 1..nx along x; rows (f = 2..ny, k, b)
    nt = (nx + 63) / 64 * 4;
    rows = (long)(ny - 1) * nz * nb;
  }
}

// tile t -> (face f, cross index tr, k, b) of lane column col; ok = exists
__device__ __forceinline__ bool tile_face(const FaceArgs &a, long t, int col, int &f, int &tr, int &k, int &b) {
  int nt;
  long rows;
  face_space(a.axis, a.nx, a.ny, a.nz, a.nb, nt, rows);
  // 32-bit index arithmetic: the host keeps the tile count below 2^31
  const unsigned tu = (unsigned)t, ntu = (unsigned)nt;
  const unsigned row = tu / ntu;
  const int x = (int)(tu - row * ntu) * 16 + col;
  const unsigned ry = (unsigned)((a.axis == 0) ? a.ny : a.ny - 1), kb = row / ry;
  const unsigned kk = kb / (unsigned)a.nz;
  k = (int)(kb - kk * (unsigned)a.nz) + 1;
  b = (int)kk;
  if (a.axis == 0) {
    f = x + 2;
    tr = (int)(row - kb * ry) + 1;
    return f <= a.nx;
  } else {
    tr = x + 1;
    f = (int)(row - kb * ry) + 2;
    return tr <= a.nx;
  }
}

// face index in the per-axis face arrays (u layout / v layout, per bin)
__device__ __forceinline__ size_t face_index(const FaceArgs &a, int f, int tr, int k, int b) {
  if (a.axis == 0)
    return (size_t)f + (size_t)(a.nx + 3) * ((size_t)(tr - 1) + (size_t)a.ny * ((size_t)(k - 1) + (size_t)a.nz * b));
  return (size_t)(tr - 1) + (size_t)a.nx * ((size_t)f + (size_t)(a.ny + 3) * ((size_t)(k - 1) + (size_t)a.nz * b));
}

// ---------------------------------------------------------------------------
// inference: dPhi of every interior face of one axis (others stay zero)
// ---------------------------------------------------------------------------
// Activity, pass 1: one bit per cell, |c| above max(smin, 1e-30), in words of
// 64 cells per row (j, k, b): cell i is bit (i-1) & 63 of word (i-1) / 64 of
// the row's nw words. Each wave takes 16 consecutive words (one cell per
// lane; a ballot makes the word). The host keeps the word count below 2^31.
__global__ void __launch_bounds__(256) k_live_bits(FaceArgs a, unsigned long long *__restrict__ bits) {
  const int l = threadIdx.x & 63;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int nw = (nx + 63) / 64;
  const unsigned nwords = (unsigned)nw * (unsigned)ny * (unsigned)nz * (unsigned)a.nb;
  const unsigned w0 = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 16u;
  if (w0 >= nwords) return;
  const float thr = fmaxf(a.smin_dev ? *a.smin_dev : a.smin, kSmallest);
  const unsigned row0 = w0 / (unsigned)nw;
  int xw = (int)(w0 - row0 * (unsigned)nw);
  int kb = (int)(row0 / (unsigned)ny);
  int j = (int)(row0 - (unsigned)kb * (unsigned)ny) + 1;
  size_t cb = cix(a, 1, j, kb % nz + 1, kb / nz);
  float v[16];
  unsigned okb = 0;
#pragma unroll
  for (int it = 0; it < 16; ++it) {  // all loads first (cells past the end read element 0)
    const int x = xw * 64 + l;
    const bool ok = w0 + it < nwords && x < nx;
    v[it] = a.c[ok ? cb + x : 0];
    okb |= (unsigned)ok << it;
    if (++xw == nw) {  // next row
      xw = 0;
      if (++j > ny) {
        j = 1;
        ++kb;
      }
      cb = cix(a, 1, j, kb % nz + 1, kb / nz);
    }
  }
#pragma unroll
  for (int it = 0; it < 16; ++it) {
    const unsigned long long m = __ballot(((okb >> it) & 1) && fabsf(v[it]) > thr);
    if (l == 0 && w0 + it < nwords) bits[w0 + it] = m;
  }
}

// Activity, pass 2 (one axis). A face is live when one of its two cells is,
// so an x tile T of row (j, k, b) (faces 16T+2 .. 16T+17) is active when one
// of the cells 16T+1 .. 16T+17 of the row is live, a y tile T of face row
// (f, k, b) when one of the cells 16T+1 .. 16T+16 of rows f-1 or f is (cells
// past nx have no bit set). The pass writes one flag per tile and appends the
// active tiles to the list. A workgroup takes kTilesPerThread x 256
// consecutive tiles and makes one atomic for all of them (the order of the
// list does not matter, every face is computed from its own stencil). Faces
// of inactive tiles get no correction: the apply pass reads their flag, not
// their dPhi. The host keeps the tile count below 2^31.
constexpr int kTilesPerThread = 8;
__device__ __forceinline__ bool tile_live(int axis, int nt, int nw, int ny,
                                          const unsigned long long *__restrict__ bits, unsigned t) {
  const unsigned row = t / (unsigned)nt;
  const int T = (int)(t - row * (unsigned)nt);
  const int w = T >> 2, o = 16 * (T & 3);
  if (w >= nw) return false;
  if (axis == 0) {  // bits 16T .. 16T+16 of cell row `row`
    const unsigned long long *r = bits + (size_t)row * nw;
    const unsigned long long lo = r[w] >> o;
    const unsigned long long hi = (o == 48 && w + 1 < nw) ? (r[w + 1] & 1ull) << 16 : 0ull;
    return ((lo | hi) & 0x1FFFFull) != 0;
  }
  // face row (f - 2) + (ny - 1) kb: cell rows f - 2 and f - 1 (0-based) of level kb
  const unsigned kb = row / (unsigned)(ny - 1);
  const unsigned f2 = row - kb * (unsigned)(ny - 1);
  const unsigned long long *r = bits + ((size_t)f2 + (size_t)ny * kb) * nw;
  return (((r[w] | r[w + nw]) >> o) & 0xFFFFull) != 0;
}

__global__ void __launch_bounds__(256) k_tile_flags(int axis, int nx, int ny, int nz, int nb,
                                                    const unsigned long long *__restrict__ bits,
                                                    unsigned char *__restrict__ flag, int *__restrict__ list,
                                                    unsigned *__restrict__ count) {
  __shared__ unsigned wcount[4], wbase[4];
  int nt;
  long rows;
  face_space(axis, nx, ny, nz, nb, nt, rows);
  const unsigned ntile = (unsigned)nt * (unsigned)rows;
  const int nw = (nx + 63) / 64;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned tb = blockIdx.x * (256u * kTilesPerThread) + threadIdx.x;
  unsigned long long m[kTilesPerThread];
  unsigned livebits = 0, n = 0;
#pragma unroll
  for (int it = 0; it < kTilesPerThread; ++it) {
    const unsigned t = tb + 256u * it;
    bool live = false;
    if (t < ntile) {
      live = tile_live(axis, nt, nw, ny, bits, t);
      flag[t] = (unsigned char)live;
    }
    m[it] = __ballot(live);
    livebits |= (unsigned)live << it;
    n += (unsigned)__popcll(m[it]);
  }
  // one atomic per workgroup; waves take consecutive ranges in wave order
  if (l == 0) wcount[wv] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned tot = 0;
    for (int q = 0; q < 4; ++q) {
      wbase[q] = tot;
      tot += wcount[q];
    }
    const unsigned b0 = tot ? atomicAdd(count, tot) : 0u;
    for (int q = 0; q < 4; ++q) wbase[q] += b0;
  }
  __syncthreads();
  unsigned off = wbase[wv];
  const unsigned long long below = (1ull << l) - 1;
#pragma unroll
  for (int it = 0; it < kTilesPerThread; ++it) {
    if ((livebits >> it) & 1) list[off + __popcll(m[it] & below)] = (int)(tb + 256u * it);
    off += (unsigned)__popcll(m[it]);
  }
}

// inference: dPhi of every face of the active tiles
// bad[0] is set when the network gives a non-finite output for an active
// face (the clamps below would otherwise turn it into a finite value)
__global__ void __launch_bounds__(256) k_correct_faces(FaceArgs a, const PackedWeights *__restrict__ w,
                                                       float *__restrict__ dphi, const int *__restrict__ list,
                                                       const unsigned *__restrict__ count,
                                                       unsigned *__restrict__ bad) {
  const int l = threadIdx.x & 63, col = l & 15;
  const long na = *count;
  const long wave = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long nwave = (long)gridDim.x * (blockDim.x >> 6);
  bf16x8 w1t[4], w2t[4][2];
  load_weights(w, w1t, w2t);
  const float smin = a.smin_dev ? *a.smin_dev : a.smin;
  for (long q = wave; q < na; q += nwave) {
    const long t = list[q];
    int f, tr, k, b;
    const bool ok = tile_face(a, t, col, f, tr, k, b);
    FaceGeo geo;
    if (ok) geo = face_geo(a, f, tr, k, b);
    float S = 0.0f;
    const bf16x8 xb = face_features(a, geo, ok, k, b, S);
    const bool act = ok && S > smin;
    float out = 0.0f;
    if (__ballot(act)) {  // uniform: skip the network on tiles without mass
      TileAct ta;
      tile_forward(w, w1t, w2t, xb, ta);
      const float y = ta.y;
      if (__ballot(act && !isfinite(y)) && l == 0) atomicOr(bad, 1u);
      if (act) {
        float F = S * geo.dxp * kOutScale * y;  // canonical direction (upwind -> downwind)
        // positivity: a face may move at most a quarter of its donor cell
        const int ax = a.axis;
        const int iD = (geo.sgn > 0.0f) ? geo.iR : geo.iL, jD = (geo.sgn > 0.0f) ? geo.jR : geo.jL;
        const float cu = a.c[cix(a, geo.iU, geo.jU, k, b)], cd = a.c[cix(a, iD, jD, k, b)];
        const float du = 1.0f / a.idxb[ax][ax == 0 ? geo.iU : geo.jU];
        const float dd = 1.0f / a.idxb[ax][ax == 0 ? iD : jD];
        F = fminf(F, 0.25f * fmaxf(cu, 0.0f) * du);
        F = fmaxf(F, -0.25f * fmaxf(cd, 0.0f) * dd);
        out = a.gain * geo.sgn * F;
      }
    }
    if (ok && l < 16) dphi[face_index(a, f, tr, k, b)] = out;
  }
}

__global__ void k_zero_counts(unsigned *__restrict__ count, int n) {
  if (threadIdx.x < n) count[threadIdx.x] = 0u;
}

// correction of one face, limited to lim |plain face flux| (lim <= 0: off);
// zero for a face outside every active tile
__device__ __forceinline__ float limited(const float *__restrict__ d, const float *__restrict__ phi, size_t i,
                                         float lim, bool on) {
  if (!on) return 0.0f;
  const float v = d[i];
  if (lim <= 0.0f) return v;
  const float m = lim * fabsf(phi[i]);
  return fminf(fmaxf(v, -m), m);
}
// Tile flags a 64-cell segment sx of cell row (j, kb) needs, kb = (k - 1) +
// nz b, as words: the 4 x tiles of its faces 64 sx + 2 .. 64 sx + 65 (one
// 32-bit word, tile rows being whole words), the x tile before them (face
// 64 sx + 1), and the 4 y tiles of the face rows j and j + 1 (zero where the
// row or word does not exist).
struct SegFlags {
  unsigned wx, wp, wy0, wy1;
};
__device__ __forceinline__ SegFlags seg_flags(const unsigned char *__restrict__ flx,
                                              const unsigned char *__restrict__ fly, int ntx, int nty, int ny,
                                              int j, int kb, int sx) {
  const unsigned *__restrict__ wfx = reinterpret_cast<const unsigned *>(flx);
  const unsigned *__restrict__ wfy = reinterpret_cast<const unsigned *>(fly);
  const size_t rx = (size_t)(j - 1) + (size_t)ny * kb;  // x-face row of these cells
  SegFlags f;
  f.wx = (sx < (ntx >> 2)) ? wfx[rx * (ntx >> 2) + sx] : 0u;
  f.wp = (sx > 0) ? flx[rx * ntx + 4 * sx - 1] : 0u;
  f.wy0 = (j >= 2) ? wfy[((size_t)(j - 2) + (size_t)(ny - 1) * kb) * (nty >> 2) + sx] : 0u;
  f.wy1 = (j + 1 <= ny) ? wfy[((size_t)(j - 1) + (size_t)(ny - 1) * kb) * (nty >> 2) + sx] : 0u;
  return f;
}

// List of the 64-cell row segments with a face in an active tile (one thread
// per segment, kTilesPerThread segments per thread, one atomic per
// workgroup). nsegs = nseg * ny * nz * nb < 2^31.
__global__ void __launch_bounds__(256) k_touched_segments(const unsigned char *__restrict__ flx,
                                                          const unsigned char *__restrict__ fly, int nx, int ny,
                                                          int nz, int nb, int nseg, int *__restrict__ list,
                                                          unsigned *__restrict__ count) {
  __shared__ unsigned wcount[4], wbase[4];
  int ntx, nty;
  long rows_unused;
  face_space(0, nx, ny, nz, nb, ntx, rows_unused);
  face_space(1, nx, ny, nz, nb, nty, rows_unused);
  const unsigned nsegs = (unsigned)nseg * (unsigned)ny * (unsigned)nz * (unsigned)nb;
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned gb = blockIdx.x * (256u * kTilesPerThread) + threadIdx.x;
  unsigned long long m[kTilesPerThread];
  unsigned tbits = 0, n = 0;
#pragma unroll
  for (int it = 0; it < kTilesPerThread; ++it) {
    const unsigned g = gb + 256u * it;
    bool t = false;
    if (g < nsegs) {
      const unsigned row = g / (unsigned)nseg;  // (j - 1) + ny kb
      const int sx = (int)(g - row * (unsigned)nseg);
      const unsigned kb = row / (unsigned)ny;
      const int j = (int)(row - kb * (unsigned)ny) + 1;
      const SegFlags f = seg_flags(flx, fly, ntx, nty, ny, j, (int)kb, sx);
      t = (f.wx | f.wp | f.wy0 | f.wy1) != 0u;
    }
    m[it] = __ballot(t);
    tbits |= (unsigned)t << it;
    n += (unsigned)__popcll(m[it]);
  }
  if (l == 0) wcount[wv] = n;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned tot = 0;
    for (int q = 0; q < 4; ++q) {
      wbase[q] = tot;
      tot += wcount[q];
    }
    const unsigned b0 = tot ? atomicAdd(count, tot) : 0u;
    for (int q = 0; q < 4; ++q) wbase[q] += b0;
  }
  __syncthreads();
  unsigned off = wbase[wv];
  const unsigned long long below = (1ull << l) - 1;
#pragma unroll
  for (int it = 0; it < kTilesPerThread; ++it) {
    if ((tbits >> it) & 1) list[off + __popcll(m[it] & below)] = (int)(gb + 256u * it);
    off += (unsigned)__popcll(m[it]);
  }
}

// Each wave takes listed segments (uniform), one cell per lane; a cell none
// of whose four faces lies in an active tile is left untouched: its
// correction is zero. Every cell is updated by one thread, so the result
// does not depend on the order of the list. When the network gave a
// non-finite correction this step (bad[0]), nothing is applied: the state
// stays the plain KT step, and bad[1] counts the fallback.
__global__ void __launch_bounds__(256) k_apply_corrections(
    float *__restrict__ c, const float *__restrict__ dx, const float *__restrict__ dy, const float *__restrict__ px,
    const float *__restrict__ py, const unsigned char *__restrict__ flx, const unsigned char *__restrict__ fly,
    float lim, int nx, int ny, int nz, int nb, int nseg, const int *__restrict__ list,
    const unsigned *__restrict__ count, const float *__restrict__ idxb_x, const float *__restrict__ idxb_y,
    unsigned *__restrict__ bad) {
  if (bad[0]) {  // uniform over the grid
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(bad + 1, 1u);
    return;
  }
  int ntx, nty;  // tiles per face row of each axis (multiples of 4)
  long rows_unused;
  face_space(0, nx, ny, nz, nb, ntx, rows_unused);
  face_space(1, nx, ny, nz, nb, nty, rows_unused);
  const unsigned na = *count;
  const unsigned nwave = gridDim.x * (blockDim.x >> 6);
  const int l = threadIdx.x & 63;
  for (unsigned q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); q < na; q += nwave) {
    const unsigned sg = (unsigned)list[q];
    const unsigned row = sg / (unsigned)nseg;  // (j - 1) + ny kb
    const int sx = (int)(sg - row * (unsigned)nseg);
    const int kb = (int)(row / (unsigned)ny);
    const int j = (int)(row - (unsigned)kb * (unsigned)ny) + 1;
    const SegFlags fl = seg_flags(flx, fly, ntx, nty, ny, j, kb, sx);
    const int i = sx * 64 + l + 1;
    if (i > nx) continue;
    // tile of face i: the word's tile ((l - 1) >> 4) (lane 0: the tile
    // before); of face i + 1: tile l >> 4; of the y faces: tile l >> 4
    const uint64_t X = (uint64_t)fl.wp | ((uint64_t)fl.wx << 8);
    const int tl0 = (l == 0) ? 0 : ((l - 1) >> 4) + 1, tl1 = (l >> 4) + 1;
    const bool ox0 = i >= 2 && ((X >> (8 * tl0)) & 1);
    const bool ox1 = i + 1 <= nx && ((X >> (8 * tl1)) & 1);
    const bool oy0 = (fl.wy0 >> (8 * (l >> 4))) & 1;
    const bool oy1 = (fl.wy1 >> (8 * (l >> 4))) & 1;
    if (!(ox0 || ox1 || oy0 || oy1)) continue;
    const int k = kb % nz + 1, b = kb / nz;
    const size_t kbs = (size_t)kb;
    const size_t fx = (size_t)i + (size_t)(nx + 3) * ((size_t)(j - 1) + (size_t)ny * kbs);
    const size_t fy = (size_t)(i - 1) + (size_t)nx * ((size_t)j + (size_t)(ny + 3) * kbs);
    const float d = (limited(dx, px, fx + 1, lim, ox1) - limited(dx, px, fx, lim, ox0)) * idxb_x[i] +
                    (limited(dy, py, fy + nx, lim, oy1) - limited(dy, py, fy, lim, oy0)) * idxb_y[j];
    const size_t ci = (size_t)(i + 1) + (size_t)state_pitch(nx) * ((size_t)(j + 1) + (size_t)(ny + 4) * ((size_t)(k + 1) + (size_t)(nz + 4) * b));
    c[ci] -= d;
  }
}

// ---------------------------------------------------------------------------
// training rows
// ---------------------------------------------------------------------------
// Append the features of every active interior face of one axis (rows
// counter[0]..; rows beyond cap are dropped), with its reference and
// denormalisation sgn S dxp.
__global__ void __launch_bounds__(256) k_collect_features(FaceArgs a, uint16_t *__restrict__ xr,
                                                          int4 *__restrict__ ref, float *__restrict__ den,
                                                          unsigned int *__restrict__ counter, long cap) {
  const int l = threadIdx.x & 63, col = l & 15, g = l >> 4;
  int nt;
  long rows;
  face_space(a.axis, a.nx, a.ny, a.nz, a.nb, nt, rows);
  const long ntile = (long)nt * rows;
  const long wave = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long nwave = (long)gridDim.x * (blockDim.x >> 6);
  for (long t = wave; t < ntile; t += nwave) {
    int f, tr, k, b;
    const bool ok = tile_face(a, t, col, f, tr, k, b);
    FaceGeo geo;
    if (ok) geo = face_geo(a, f, tr, k, b);
    float S = 0.0f;
    const bf16x8 xb = face_features(a, geo, ok, k, b, S);
    const bool act = ok && S > a.smin;
    // rank of this face among the tile's active faces (group 0 lanes decide)
    const unsigned long long m = __ballot(act && g == 0);
    if (m == 0) continue;
    const int nact = __popcll(m);
    unsigned int base = 0;
    if (l == 0) base = atomicAdd(counter, (unsigned int)nact);
    base = __shfl(base, 0, 64);
    const int rank = __popcll(m & ((1ull << col) - 1ull));
    const long row = (long)base + rank;
    if (act && row < cap) {
      *reinterpret_cast<bf16x8 *>(xr + row * 32 + 8 * g) = xb;
      if (g == 0) {
        ref[row] = make_int4(a.axis | (b << 1), f, tr, k);
        den[row] = geo.sgn * S * geo.dxp;
      }
    }
  }
}

// Canonical order of the rows of one collection: a face's key orders by bin,
// axis, k, transverse index and face index (radices nk, nt, nf above their
// largest values), so the rows, and the batches sampled from them, do not
// depend on the order in which waves appended them.
__global__ void k_row_keys(const int4 *__restrict__ ref, long m, int nk, int nt, int nf,
                           unsigned long long *__restrict__ key, unsigned int *__restrict__ idx) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int4 q = ref[i];
  const unsigned long long ba = 2ull * (unsigned long long)(q.x >> 1) + (unsigned long long)(q.x & 1);
  key[i] = ((ba * (unsigned)nk + (unsigned)q.w) * (unsigned)nt + (unsigned)q.z) * (unsigned)nf + (unsigned)q.y;
  idx[i] = (unsigned int)i;
}

// row i of the sorted order := row perm[i] (features, face, denominator)
__global__ void k_gather_rows(const uint16_t *__restrict__ xs, const int4 *__restrict__ rs,
                              const float *__restrict__ ds, const unsigned int *__restrict__ perm, long m,
                              uint16_t *__restrict__ xd, int4 *__restrict__ rd, float *__restrict__ dd) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= m) return;
  const long src = perm[i];
  const uint4 *xsrc = reinterpret_cast<const uint4 *>(xs + src * 32);
  uint4 *xdst = reinterpret_cast<uint4 *>(xd + i * 32);
#pragma unroll
  for (int q = 0; q < 4; ++q) xdst[q] = xsrc[q];
  rd[i] = rs[src];
  dd[i] = ds[src];
}

// targets of rows r0..r1: (fine flux averaged over the two fine faces of
// the coarse face - coarse flux) / (kOutScale den)
__global__ void k_collect_targets(const int4 *__restrict__ ref, const float *__restrict__ den, long r0, long r1,
                                  const float *__restrict__ cfx, const float *__restrict__ cfy,
                                  const float *__restrict__ ffx, const float *__restrict__ ffy, int nx, int ny,
                                  int nz, float *__restrict__ tgt) {
  const long r = r0 + blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r >= r1) return;
  const int4 q = ref[r];
  const int axis = q.x & 1, b = q.x >> 1, f = q.y, tr = q.z, k = q.w;
  const int fnx = 2 * nx, fny = 2 * ny;
  float pc, pf;
  if (axis == 0) {
    pc = cfx[(size_t)f + (size_t)(nx + 3) * ((size_t)(tr - 1) + (size_t)ny * ((size_t)(k - 1) + (size_t)nz * b))];
    const int F = 2 * f - 1;
    pf = 0.0f;
    for (int d = 0; d < 2; ++d) {
      const int jf = 2 * tr - 1 + d;
      pf += ffx[(size_t)F + (size_t)(fnx + 3) * ((size_t)(jf - 1) + (size_t)fny * ((size_t)(k - 1) + (size_t)nz * b))];
    }
  } else {
    pc = cfy[(size_t)(tr - 1) + (size_t)nx * ((size_t)f + (size_t)(ny + 3) * ((size_t)(k - 1) + (size_t)nz * b))];
    const int F = 2 * f - 1;
    pf = 0.0f;
    for (int d = 0; d < 2; ++d) {
      const int iff = 2 * tr - 1 + d;
      pf += ffy[(size_t)(iff - 1) + (size_t)fnx * ((size_t)F + (size_t)(fny + 3) * ((size_t)(k - 1) + (size_t)nz * b))];
    }
  }
  const float dn = den[r];
  tgt[r] = (dn != 0.0f) ? (0.5f * pf - pc) / (kOutScale * dn) : 0.0f;
}

// largest |c| of a field: per-block partial maxima, then the threshold
// rel * max in out[0] (maxima do not depend on the order)
__global__ void k_absmax(const float *__restrict__ c, size_t n, float *__restrict__ part) {
  __shared__ float red[4];
  float m = 0.0f;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(c[q]));
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

__global__ void k_threshold(const float *__restrict__ part, int nb, float rel, float *__restrict__ out) {
  float m = 0.0f;
  for (int b = threadIdx.x; b < nb; b += 64) m = fmaxf(m, part[b]);
  for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
  if (threadIdx.x == 0) out[0] = rel * m;
}

// coarse(I, J, k, b) = mean of the 2x2 fine cells (interior only)
__global__ void k_coarse_grain(const float *__restrict__ fine, float *__restrict__ coarse, int nx, int ny, int nz,
                               int nb) {
  const size_t n = (size_t)nx * ny * nz * nb;
  const int fnx = 2 * nx, fny = 2 * ny;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const int I = (int)(q % nx) + 1;
    const int J = (int)((q / nx) % ny) + 1;
    const int k = (int)((q / ((size_t)nx * ny)) % nz) + 1;
    const int b = (int)(q / ((size_t)nx * ny * nz));
    float s = 0.0f;
    for (int dj = 0; dj < 2; ++dj)
      for (int di = 0; di < 2; ++di) {
        const int i = 2 * I - 1 + di, j = 2 * J - 1 + dj;
        s += fine[(size_t)(i + 1) + (size_t)state_pitch(fnx) * ((size_t)(j + 1) + (size_t)(fny + 4) * ((size_t)(k + 1) + (size_t)(nz + 4) * b))];
      }
    coarse[(size_t)(I + 1) + (size_t)state_pitch(nx) * ((size_t)(J + 1) + (size_t)(ny + 4) * ((size_t)(k + 1) + (size_t)(nz + 4) * b))] = 0.25f * s;
  }
}

// uniform batch rows: idx[r] = hash(seed, step, r) mod count (SplitMix64)
__global__ void k_batch_rows(uint32_t *__restrict__ idx, long batch, long count, unsigned long long seed, int step) {
  const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r >= batch) return;
  unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(step * 0x100000000ll + r + 1);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  idx[r] = (uint32_t)(z % (unsigned long long)count);
}

__global__ void k_copy_f(const float *__restrict__ src, float *__restrict__ dst, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace f3d

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

int nfail(fall3d_mlp *n, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (n) n->err = buf;
  return code;
}

#define NHIPCHK(net, expr)                                                                            \
  do {                                                                                                \
    hipError_t e_ = (expr);                                                                           \
    if (e_ != hipSuccess) return nfail((net), FALL3D_ADS_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

FaceArgs face_args(fall3d_ads_ctx *c, int axis, double dt, double stime, double smin) {
  FaceArgs a;
  memset(&a, 0, sizeof(a));
  a.c = c->chi;
  a.nx = c->nx;
  a.ny = c->ny;
  a.nz = c->nz;
  a.nb = c->nb;
  a.axis = axis;
  a.u1 = c->u1;
  a.u2 = c->u2;
  a.v1 = c->v1;
  a.v2 = c->v2;
  a.w1 = c->w1;
  a.w2 = c->w2;
  a.vs = c->vs;
  a.s = (float)stime;
  a.s1 = (float)(1.0 - stime);
  a.dt = (float)dt;
  for (int d = 0; d < 3; ++d) {
    a.hdxp[d] = c->hdxp[d];
    a.idxb[d] = c->idxb[d];
  }
  a.smin = (float)smin;
  a.smin_dev = nullptr;
  a.gain = 1.0f;
  return a;
}

size_t nface(const fall3d_ads_ctx *c, int axis) {
  return axis == 0 ? (size_t)(c->nx + 3) * c->ny * c->nz * c->nb : (size_t)(c->ny + 3) * c->nx * c->nz * c->nb;
}

unsigned grid_for(long ntile) { return (unsigned)std::max<long>(1, std::min<long>((ntile + 3) / 4, 8192)); }

long ntiles(const fall3d_ads_ctx *c, int axis) {
  int nt = 0;
  long rows = 0;
  face_space(axis, c->nx, c->ny, c->nz, c->nb, nt, rows);
  return (long)nt * rows;
}

int pack(fall3d_mlp *n) {
  NHIPCHK(n, mlp_pack(n->params, n->packed, n->stream));
  return FALL3D_ADS_OK;
}

int ensure_part(fall3d_mlp *n, long rows) {
  const int waves = mlp_grad_waves(rows);
  if (waves <= n->part_waves) return FALL3D_ADS_OK;
  dfree(n->part);
  n->part = nullptr;
  n->part_waves = 0;
  NHIPCHK(n, dalloc(&n->part, (size_t)waves * (MLP_NPARAM + 1)));
  n->part_waves = waves;
  return FALL3D_ADS_OK;
}

}  // namespace

namespace f3d {
namespace host {

int learned_pre_step(fall3d_ads_ctx *c, double dt, double stime) {
  const bool rel = c->learned_smin < 0.0;
  if (rel) {  // threshold relative to the field's largest value
    hipLaunchKernelGGL(k_absmax, dim3(c->lthr_blocks), dim3(256), 0, c->stream, c->chi, c->ncell_s, c->lthr + 1);
    hipLaunchKernelGGL(k_threshold, dim3(1), dim3(64), 0, c->stream, c->lthr + 1, c->lthr_blocks,
                       (float)(-c->learned_smin), c->lthr);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, FALL3D_ADS_EHIP, "learned threshold: %s", hipGetErrorString(e));
  }
  // a fill kernel, not a memset: the step may be captured in a graph
  hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(64), 0, c->stream, c->lcount, 3);
  hipLaunchKernelGGL(k_zero_counts, dim3(1), dim3(64), 0, c->stream, c->lbad, 1);
  for (int axis = 0; axis < 2; ++axis) {
    FaceArgs a = face_args(c, axis, dt, stime, rel ? 0.0 : c->learned_smin);
    if (rel) a.smin_dev = c->lthr;
    a.gain = (float)c->learned_gain;
    const long nt = ntiles(c, axis);
    if (axis == 0) {  // live cells of the state, for both axes: 16 words of 64 cells per wave
      const long words = (long)c->ny * c->nz * c->nb * ((c->nx + 63) / 64);
      hipLaunchKernelGGL(k_live_bits, dim3((unsigned)((words + 63) / 64)), dim3(256), 0, c->stream, a, c->lbits);
    }
    hipLaunchKernelGGL(k_tile_flags, dim3((unsigned)((nt + 256 * kTilesPerThread - 1) / (256 * kTilesPerThread))),
                       dim3(256), 0, c->stream, axis, c->nx, c->ny,
                       c->nz, c->nb, c->lbits, c->lflag[axis], c->lact[axis], c->lcount + axis);
    hipLaunchKernelGGL(k_correct_faces, dim3(grid_for(nt)), dim3(256), 0, c->stream, a, c->learned->packed,
                       c->dphi[axis], c->lact[axis], c->lcount + axis, c->lbad);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, FALL3D_ADS_EHIP, "face corrections: %s", hipGetErrorString(e));
  }
  return FALL3D_ADS_OK;
}

int learned_post_step(fall3d_ads_ctx *c) {
  const int nseg = (c->nx + 63) / 64;
  const long nsegs = (long)c->ny * c->nz * c->nb * nseg;
  const float lim = (float)c->learned_flux_limit;
  hipLaunchKernelGGL(k_touched_segments, dim3((unsigned)((nsegs + 256 * kTilesPerThread - 1) / (256 * kTilesPerThread))),
                     dim3(256), 0, c->stream, c->lflag[0],
                     c->lflag[1], c->nx, c->ny, c->nz, c->nb, nseg, c->lseg, c->lcount + 2);
  const unsigned blocks = (unsigned)std::min<long>((nsegs + 3) / 4, 8192);
  hipLaunchKernelGGL(k_apply_corrections, dim3(blocks), dim3(256), 0, c->stream, c->chi, c->dphi[0], c->dphi[1],
                     lim > 0.0f ? c->fout[0] : c->dphi[0], lim > 0.0f ? c->fout[1] : c->dphi[1], c->lflag[0],
                     c->lflag[1], lim, c->nx, c->ny, c->nz, c->nb, nseg, c->lseg, c->lcount + 2, c->idxb[0],
                     c->idxb[1], c->lbad);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, FALL3D_ADS_EHIP, "apply corrections: %s", hipGetErrorString(e));
  return FALL3D_ADS_OK;
}

void learned_release(fall3d_ads_ctx *c) {
  for (int a = 0; a < 2; ++a) {
    dfree(c->fout[a]);
    dfree(c->dphi[a]);
    c->fout[a] = c->dphi[a] = nullptr;
  }
  dfree(c->lthr);
  c->lthr = nullptr;
  for (int a = 0; a < 2; ++a) {
    dfree(c->lact[a]);
    dfree(c->lflag[a]);
    c->lact[a] = nullptr;
    c->lflag[a] = nullptr;
  }
  dfree(c->lseg);
  c->lseg = nullptr;
  dfree(c->lbits);
  c->lbits = nullptr;
  dfree(c->lcount);
  c->lcount = nullptr;
  dfree(c->lbad);
  c->lbad = nullptr;
  c->learned = nullptr;
}

}  // namespace host
}  // namespace f3d

// Rows r0..r1 of n into their canonical order (k_row_keys): a radix sort of
// the face keys, then a gather through temporary buffers.
static int sort_rows(fall3d_mlp *n, fall3d_ads_ctx *coarse, long r0, long r1, hipStream_t st) {
  const long m = r1 - r0;
  if (m < 2) return FALL3D_ADS_OK;
  if (m > (long)INT_MAX) return nfail(n, FALL3D_ADS_EINVAL, "too many rows in one collection (%ld)", m);
  const int nk = coarse->nz + 1, nt = std::max(coarse->nx, coarse->ny) + 1, nf = nt;
  const double kmax = 2.0 * coarse->nb * (double)nk * nt * nf;
  if (kmax >= 18446744073709551615.0) return nfail(n, FALL3D_ADS_EINVAL, "face keys exceed 64 bits");
  int nbits = 1;
  while (nbits < 64 && std::ldexp(1.0, nbits) < kmax) ++nbits;
  unsigned long long *key = nullptr, *key2 = nullptr;
  unsigned int *idx = nullptr, *idx2 = nullptr;
  uint16_t *xt = nullptr;
  int4 *rt = nullptr;
  float *dtmp = nullptr;
  void *tmp = nullptr;
  size_t tbytes = 0;
  hipError_t e = hipSuccess;
  auto take = [&](hipError_t r) {
    if (e == hipSuccess) e = r;
  };
  take(hipMalloc(&key, m * sizeof(*key)));
  take(hipMalloc(&key2, m * sizeof(*key2)));
  take(hipMalloc(&idx, m * sizeof(*idx)));
  take(hipMalloc(&idx2, m * sizeof(*idx2)));
  take(hipMalloc(&xt, m * 32 * sizeof(*xt)));
  take(hipMalloc(&rt, m * sizeof(*rt)));
  take(hipMalloc(&dtmp, m * sizeof(*dtmp)));
  if (e == hipSuccess) {
    const unsigned g = (unsigned)((m + 255) / 256);
    hipLaunchKernelGGL(k_row_keys, dim3(g), dim3(256), 0, st, n->ref + r0, m, nk, nt, nf, key, idx);
    take(hipGetLastError());
    take(hipcub::DeviceRadixSort::SortPairs(nullptr, tbytes, key, key2, idx, idx2, (int)m, 0, nbits, st));
    take(hipMalloc(&tmp, tbytes > 0 ? tbytes : 1));
    take(hipcub::DeviceRadixSort::SortPairs(tmp, tbytes, key, key2, idx, idx2, (int)m, 0, nbits, st));
    if (e == hipSuccess) {
      hipLaunchKernelGGL(k_gather_rows, dim3(g), dim3(256), 0, st, n->x + r0 * 32, n->ref + r0, n->den + r0, idx2, m,
                         xt, rt, dtmp);
      take(hipGetLastError());
      take(hipMemcpyAsync(n->x + r0 * 32, xt, m * 32 * sizeof(*xt), hipMemcpyDeviceToDevice, st));
      take(hipMemcpyAsync(n->ref + r0, rt, m * sizeof(*rt), hipMemcpyDeviceToDevice, st));
      take(hipMemcpyAsync(n->den + r0, dtmp, m * sizeof(*dtmp), hipMemcpyDeviceToDevice, st));
      take(hipStreamSynchronize(st));
    }
  }
  for (void *p : {(void *)key, (void *)key2, (void *)idx, (void *)idx2, (void *)xt, (void *)rt, (void *)dtmp, tmp})
    if (p) (void)hipFree(p);
  if (e != hipSuccess) return nfail(n, FALL3D_ADS_EHIP, "row sort: %s", hipGetErrorString(e));
  return FALL3D_ADS_OK;
}

extern "C" {

int fall3d_mlp_create(fall3d_mlp **pnet, long capacity) {
  if (!pnet) return FALL3D_ADS_EINVAL;
  auto *n = new fall3d_mlp();
  *pnet = n;
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess)
    return nfail(n, FALL3D_ADS_EHIP, "stream creation failed");
  NHIPCHK(n, dalloc(&n->params, MLP_NPARAM));
  NHIPCHK(n, dalloc(&n->m, MLP_NPARAM));
  NHIPCHK(n, dalloc(&n->v, MLP_NPARAM));
  NHIPCHK(n, dalloc(&n->grad, MLP_NPARAM + 1));
  NHIPCHK(n, dalloc(&n->packed, 1));
  NHIPCHK(n, dalloc(&n->dcount, 1));
  NHIPCHK(n, hipMemsetAsync(n->params, 0, sizeof(float) * MLP_NPARAM, n->stream));
  NHIPCHK(n, hipMemsetAsync(n->m, 0, sizeof(float) * MLP_NPARAM, n->stream));
  NHIPCHK(n, hipMemsetAsync(n->v, 0, sizeof(float) * MLP_NPARAM, n->stream));
  NHIPCHK(n, hipMemsetAsync(n->dcount, 0, sizeof(unsigned int), n->stream));
  if (capacity > 0) {
    if (capacity > (1l << 31)) return nfail(n, FALL3D_ADS_EINVAL, "capacity above 2^31 rows");
    n->cap = capacity;
    NHIPCHK(n, dalloc(&n->x, (size_t)capacity * 32));
    NHIPCHK(n, dalloc(&n->t, (size_t)capacity));
    NHIPCHK(n, dalloc(&n->ref, (size_t)capacity));
    NHIPCHK(n, dalloc(&n->den, (size_t)capacity));
  }
  int rc = pack(n);
  if (rc) return rc;
  NHIPCHK(n, hipStreamSynchronize(n->stream));
  return FALL3D_ADS_OK;
}

int fall3d_mlp_release(fall3d_mlp *n) {
  if (!n) return FALL3D_ADS_OK;
  if (n->stream) (void)hipStreamSynchronize(n->stream);
  void *p[] = {n->params, n->m, n->v, n->grad, n->part, n->packed, n->x, n->t, n->dcount, n->ref, n->den, n->bidx,
               n->loss};
  for (void *q : p) dfree(q);
  if (n->stream) (void)hipStreamDestroy(n->stream);
  delete n;
  return FALL3D_ADS_OK;
}

const char *fall3d_mlp_last_error(const fall3d_mlp *n) { return n ? n->err.c_str() : "null network"; }

int fall3d_mlp_nparam(void) { return MLP_NPARAM; }

int fall3d_mlp_set_params(fall3d_mlp *n, const float *params) {
  if (!n || !params) return FALL3D_ADS_EINVAL;
  NHIPCHK(n, hipMemcpyAsync(n->params, params, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, n->stream));
  NHIPCHK(n, hipMemsetAsync(n->m, 0, sizeof(float) * MLP_NPARAM, n->stream));
  NHIPCHK(n, hipMemsetAsync(n->v, 0, sizeof(float) * MLP_NPARAM, n->stream));
  n->adam_t = 0;
  int rc = pack(n);
  if (rc) return rc;
  NHIPCHK(n, hipStreamSynchronize(n->stream));
  return FALL3D_ADS_OK;
}

int fall3d_mlp_get_params(fall3d_mlp *n, float *params) {
  if (!n || !params) return FALL3D_ADS_EINVAL;
  NHIPCHK(n, hipMemcpyAsync(params, n->params, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, n->stream));
  NHIPCHK(n, hipStreamSynchronize(n->stream));
  return FALL3D_ADS_OK;
}

int fall3d_mlp_forward(fall3d_mlp *n, const uint16_t *x, long rows, float *y) {
  if (!n || (!x && rows > 0) || (!y && rows > 0) || rows < 0) return FALL3D_ADS_EINVAL;
  if (rows == 0) return FALL3D_ADS_OK;
  uint16_t *dx = nullptr;
  float *dy = nullptr;
  NHIPCHK(n, dalloc(&dx, (size_t)rows * 32));
  hipError_t e = dalloc(&dy, (size_t)rows);
  if (e == hipSuccess) e = hipMemcpyAsync(dx, x, sizeof(uint16_t) * 32 * rows, hipMemcpyHostToDevice, n->stream);
  if (e == hipSuccess) e = mlp_forward(n->packed, dx, rows, dy, n->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(y, dy, sizeof(float) * rows, hipMemcpyDeviceToHost, n->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
  dfree(dx);
  dfree(dy);
  if (e != hipSuccess) return nfail(n, FALL3D_ADS_EHIP, "forward: %s", hipGetErrorString(e));
  return FALL3D_ADS_OK;
}

int fall3d_mlp_gradient(fall3d_mlp *n, const uint16_t *x, const float *t, long rows, float *grad, double *sse) {
  if (!n || !x || !t || !grad || rows <= 0) return FALL3D_ADS_EINVAL;
  int rc = ensure_part(n, rows);
  if (rc) return rc;
  uint16_t *dx = nullptr;
  float *dt = nullptr;
  NHIPCHK(n, dalloc(&dx, (size_t)rows * 32));
  hipError_t e = dalloc(&dt, (size_t)rows);
  std::vector<float> g(MLP_NPARAM + 1);
  if (e == hipSuccess) e = hipMemcpyAsync(dx, x, sizeof(uint16_t) * 32 * rows, hipMemcpyHostToDevice, n->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dt, t, sizeof(float) * rows, hipMemcpyHostToDevice, n->stream);
  if (e == hipSuccess)
    e = mlp_grad(n->packed, dx, dt, nullptr, rows, 1.0f / (float)rows, n->part, mlp_grad_waves(rows), n->grad,
                 n->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(g.data(), n->grad, sizeof(float) * (MLP_NPARAM + 1), hipMemcpyDeviceToHost, n->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
  dfree(dx);
  dfree(dt);
  if (e != hipSuccess) return nfail(n, FALL3D_ADS_EHIP, "gradient: %s", hipGetErrorString(e));
  memcpy(grad, g.data(), sizeof(float) * MLP_NPARAM);
  if (sse) *sse = g[MLP_NPARAM];
  return FALL3D_ADS_OK;
}

int fall3d_mlp_data_append(fall3d_mlp *n, const uint16_t *x, const float *t, long rows) {
  if (!n || rows < 0 || (rows > 0 && (!x || !t))) return FALL3D_ADS_EINVAL;
  if (n->count + rows > n->cap) return nfail(n, FALL3D_ADS_EINVAL, "training set full (%ld of %ld rows)", n->count, n->cap);
  NHIPCHK(n, hipMemcpyAsync(n->x + n->count * 32, x, sizeof(uint16_t) * 32 * rows, hipMemcpyHostToDevice, n->stream));
  NHIPCHK(n, hipMemcpyAsync(n->t + n->count, t, sizeof(float) * rows, hipMemcpyHostToDevice, n->stream));
  NHIPCHK(n, hipStreamSynchronize(n->stream));
  n->count += rows;
  const unsigned int c32 = (unsigned int)n->count;
  NHIPCHK(n, hipMemcpy(n->dcount, &c32, sizeof(c32), hipMemcpyHostToDevice));
  return FALL3D_ADS_OK;
}

int fall3d_mlp_data_count(fall3d_mlp *n, long *rows) {
  if (!n || !rows) return FALL3D_ADS_EINVAL;
  *rows = n->count;
  return FALL3D_ADS_OK;
}

int fall3d_mlp_data_get(fall3d_mlp *n, long r0, long rows, uint16_t *x, float *t) {
  if (!n || r0 < 0 || rows < 0 || r0 + rows > n->count) return FALL3D_ADS_EINVAL;
  if (x) NHIPCHK(n, hipMemcpyAsync(x, n->x + r0 * 32, sizeof(uint16_t) * 32 * rows, hipMemcpyDeviceToHost, n->stream));
  if (t) NHIPCHK(n, hipMemcpyAsync(t, n->t + r0, sizeof(float) * rows, hipMemcpyDeviceToHost, n->stream));
  NHIPCHK(n, hipStreamSynchronize(n->stream));
  return FALL3D_ADS_OK;
}

int fall3d_mlp_data_refs(fall3d_mlp *n, long r0, long rows, int32_t *ref, float *den) {
  if (!n || r0 < 0 || rows < 0 || r0 + rows > n->count) return FALL3D_ADS_EINVAL;
  if (ref) NHIPCHK(n, hipMemcpyAsync(ref, n->ref + r0, sizeof(int4) * rows, hipMemcpyDeviceToHost, n->stream));
  if (den) NHIPCHK(n, hipMemcpyAsync(den, n->den + r0, sizeof(float) * rows, hipMemcpyDeviceToHost, n->stream));
  NHIPCHK(n, hipStreamSynchronize(n->stream));
  return FALL3D_ADS_OK;
}

int fall3d_mlp_data_clear(fall3d_mlp *n) {
  if (!n) return FALL3D_ADS_EINVAL;
  n->count = 0;
  NHIPCHK(n, hipMemsetAsync(n->dcount, 0, sizeof(unsigned int), n->stream));
  NHIPCHK(n, hipStreamSynchronize(n->stream));
  return FALL3D_ADS_OK;
}

int fall3d_mlp_train(fall3d_mlp *n, int steps, int batch, float lr, unsigned long long seed, float *loss) {
  if (!n || steps < 0 || batch < 1) return FALL3D_ADS_EINVAL;
  if (n->count < 1) return nfail(n, FALL3D_ADS_ESTATE, "no training rows");
  int rc = ensure_part(n, batch);
  if (rc) return rc;
  if (n->bidx_cap < batch) {
    dfree(n->bidx);
    n->bidx = nullptr;
    NHIPCHK(n, dalloc(&n->bidx, (size_t)batch));
    n->bidx_cap = batch;
  }
  if (n->loss_cap < steps) {
    dfree(n->loss);
    n->loss = nullptr;
    NHIPCHK(n, dalloc(&n->loss, (size_t)std::max(steps, 1)));
    n->loss_cap = steps;
  }
  const int waves = mlp_grad_waves(batch);
  for (int s = 0; s < steps; ++s) {
    hipLaunchKernelGGL(k_batch_rows, dim3((batch + 255) / 256), dim3(256), 0, n->stream, n->bidx, (long)batch,
                       n->count, seed, s + n->adam_t);
    NHIPCHK(n, mlp_grad(n->packed, n->x, n->t, n->bidx, batch, 1.0f / (float)batch, n->part, waves, n->grad,
                        n->stream));
    NHIPCHK(n, mlp_adam(n->params, n->grad, n->m, n->v, n->adam_t + s + 1, lr, 0.9f, 0.999f, 1e-8f, n->stream));
    int r = pack(n);
    if (r) return r;
    hipLaunchKernelGGL(k_copy_f, dim3(1), dim3(64), 0, n->stream, n->grad + MLP_NPARAM, n->loss + s, 1);
  }
  n->adam_t += steps;
  if (loss && steps > 0) {
    NHIPCHK(n, hipMemcpyAsync(loss, n->loss, sizeof(float) * steps, hipMemcpyDeviceToHost, n->stream));
    NHIPCHK(n, hipStreamSynchronize(n->stream));
    for (int s = 0; s < steps; ++s) loss[s] /= (float)batch;  // mean squared error of each batch
  }
  NHIPCHK(n, hipStreamSynchronize(n->stream));
  return FALL3D_ADS_OK;
}

// ---- context side --------------------------------------------------------------

int fall3d_ads_set_flux_output(fall3d_ads_ctx *c, int on) {
  if (!c) return FALL3D_ADS_EINVAL;
  if (on) {
    if (c->limiter != 1 || c->comp || c->tm != 2)
      return fail(c, FALL3D_ADS_EINVAL, "face-flux output needs minmod, RK4 and the FP32 state");
    for (int a = 0; a < 2; ++a) {
      if (!c->fout[a]) HIPCHK(c, dalloc(&c->fout[a], nface(c, a)));
      HIPCHK(c, hipMemsetAsync(c->fout[a], 0, sizeof(float) * nface(c, a), c->stream));
    }
  }
  c->fout_on = on != 0;
  c->version++;  // a captured step graph must be re-recorded
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FALL3D_ADS_OK;
}

int fall3d_ads_get_face_flux(fall3d_ads_ctx *c, int axis, float *out) {
  if (!c || !out || axis < 0 || axis > 1) return FALL3D_ADS_EINVAL;
  if (!c->fout[axis]) return fail(c, FALL3D_ADS_ESTATE, "face-flux output was never enabled");
  HIPCHK(c, hipMemcpyAsync(out, c->fout[axis], sizeof(float) * nface(c, axis), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FALL3D_ADS_OK;
}

int fall3d_ads_set_learned(fall3d_ads_ctx *c, fall3d_mlp *net, double smin, double gain, double flux_limit) {
  if (!c) return FALL3D_ADS_EINVAL;
  if (net) {
    if (c->comp) return fail(c, FALL3D_ADS_EINVAL, "the learned correction runs on the FP32 state only");
    if (flux_limit > 0.0 && (c->limiter != 1 || c->tm != 2))
      return fail(c, FALL3D_ADS_EINVAL, "the flux-relative limit needs minmod and RK4");
    if (ntiles(c, 0) > INT_MAX || ntiles(c, 1) > INT_MAX || (long)c->ny * c->nz * c->nb * ((c->nx + 63) / 64) > INT_MAX)
      return fail(c, FALL3D_ADS_EINVAL, "domain too large for the learned correction's tile list");
    if (!c->lcount) HIPCHK(c, dalloc(&c->lcount, 3));
    if (!c->lbad) HIPCHK(c, dalloc(&c->lbad, 2));
    HIPCHK(c, hipMemsetAsync(c->lbad, 0, 2 * sizeof(unsigned), c->stream));
    if (!c->lseg) HIPCHK(c, dalloc(&c->lseg, (size_t)c->ny * c->nz * c->nb * ((c->nx + 63) / 64)));
    if (!c->lbits) HIPCHK(c, dalloc(&c->lbits, (size_t)c->ny * c->nz * c->nb * ((c->nx + 63) / 64)));
    for (int a = 0; a < 2; ++a) {
      if (!c->dphi[a]) HIPCHK(c, dalloc(&c->dphi[a], nface(c, a)));
      if (!c->lact[a]) HIPCHK(c, dalloc(&c->lact[a], (size_t)ntiles(c, a)));
      if (!c->lflag[a]) HIPCHK(c, dalloc(&c->lflag[a], (size_t)ntiles(c, a)));
      HIPCHK(c, hipMemsetAsync(c->dphi[a], 0, sizeof(float) * nface(c, a), c->stream));
      if (flux_limit > 0.0 && !c->fout[a]) {
        HIPCHK(c, dalloc(&c->fout[a], nface(c, a)));
        HIPCHK(c, hipMemsetAsync(c->fout[a], 0, sizeof(float) * nface(c, a), c->stream));
      }
    }
    if (!c->lthr) {
      c->lthr_blocks = 1024;
      HIPCHK(c, dalloc(&c->lthr, (size_t)c->lthr_blocks + 1));
    }
  }
  c->learned = net;
  c->learned_smin = smin;
  c->learned_gain = gain;
  c->learned_flux_limit = net ? flux_limit : 0.0;
  c->version++;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FALL3D_ADS_OK;
}

int fall3d_ads_learned_fallbacks(fall3d_ads_ctx *c, long long *steps) {
  if (!c || !steps) return FALL3D_ADS_EINVAL;
  *steps = 0;
  if (!c->lbad) return FALL3D_ADS_OK;
  unsigned v[2] = {0u, 0u};
  HIPCHK(c, hipMemcpyAsync(v, c->lbad, sizeof(v), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *steps = (long long)v[1];
  return FALL3D_ADS_OK;
}

int fall3d_ads_learned_collect(fall3d_ads_ctx *fine, fall3d_ads_ctx *coarse, fall3d_mlp *n, double dt,
                               double stime, int iiter, double smin, long *added) {
  if (!fine || !coarse || !n) return FALL3D_ADS_EINVAL;
  if (fine->nx != 2 * coarse->nx || fine->ny != 2 * coarse->ny || fine->nz != coarse->nz || fine->nb != coarse->nb)
    return nfail(n, FALL3D_ADS_EINVAL, "fine grid must be the 2x horizontal refinement of the coarse grid");
  if (n->cap == 0) return nfail(n, FALL3D_ADS_ESTATE, "network created without training capacity");
  if (is_fp64(fine) || is_fp64(coarse)) return nfail(n, FALL3D_ADS_EINVAL, "training rows come from FP32 states");
  if (ntiles(coarse, 0) > INT_MAX || ntiles(coarse, 1) > INT_MAX)
    return nfail(n, FALL3D_ADS_EINVAL, "coarse domain too large for the face tiles");
  int rc;
  hipStream_t st = coarse->stream;
  if (fine->stream != st) HIPCHK(fine, hipStreamSynchronize(fine->stream));
  // 1. coarse state := coarse-grained fine state
  {
    const size_t nc = (size_t)coarse->nx * coarse->ny * coarse->nz * coarse->nb;
    hipLaunchKernelGGL(k_coarse_grain, dim3((unsigned)std::min<size_t>((nc + 255) / 256, 65536)), dim3(256), 0, st,
                       fine->chi, coarse->chi, coarse->nx, coarse->ny, coarse->nz, coarse->nb);
    HIPCHK(coarse, hipGetLastError());
  }
  // 2. features of the active coarse faces at the start of the step
  const long r0 = n->count;
  const unsigned int r0u = (unsigned int)r0;
  NHIPCHK(n, hipMemcpyAsync(n->dcount, &r0u, sizeof(r0u), hipMemcpyHostToDevice, st));
  for (int axis = 0; axis < 2; ++axis) {
    FaceArgs a = face_args(coarse, axis, dt, stime, smin);
    hipLaunchKernelGGL(k_collect_features, dim3(grid_for(ntiles(coarse, axis))), dim3(256), 0, st, a, n->x, n->ref,
                       n->den, n->dcount, n->cap);
    HIPCHK(coarse, hipGetLastError());
  }
  unsigned int cnt = 0;
  HIPCHK(coarse, hipMemcpyAsync(&cnt, n->dcount, sizeof(cnt), hipMemcpyDeviceToHost, st));
  HIPCHK(coarse, hipStreamSynchronize(st));
  const long r1 = std::min<long>((long)cnt, n->cap);
  if ((rc = sort_rows(n, coarse, r0, r1, st))) return rc;
  // 3. one plain coarse step and two fine steps with face-flux output
  fall3d_mlp *keep_c = coarse->learned, *keep_f = fine->learned;
  coarse->learned = fine->learned = nullptr;
  struct Restore {
    fall3d_ads_ctx *c, *f;
    fall3d_mlp *kc, *kf;
    ~Restore() {
      c->learned = kc;
      f->learned = kf;
    }
  } restore{coarse, fine, keep_c, keep_f};
  if ((rc = fall3d_ads_set_flux_output(coarse, 1))) return rc;
  if ((rc = fall3d_ads_step(coarse, dt, stime, iiter))) return rc;
  if ((rc = fall3d_ads_set_flux_output(fine, 1))) return rc;
  if ((rc = fall3d_ads_step(fine, 0.5 * dt, stime, 2 * iiter))) return rc;
  if ((rc = fall3d_ads_step(fine, 0.5 * dt, stime, 2 * iiter + 1))) return rc;
  HIPCHK(fine, hipStreamSynchronize(fine->stream));
  // 4. targets
  if (r1 > r0) {
    hipLaunchKernelGGL(k_collect_targets, dim3((unsigned)((r1 - r0 + 255) / 256)), dim3(256), 0, st, n->ref, n->den,
                       r0, r1, coarse->fout[0], coarse->fout[1], fine->fout[0], fine->fout[1], coarse->nx,
                       coarse->ny, coarse->nz, n->t);
    HIPCHK(coarse, hipGetLastError());
  }
  if ((rc = fall3d_ads_set_flux_output(coarse, 0))) return rc;
  if ((rc = fall3d_ads_set_flux_output(fine, 0))) return rc;
  HIPCHK(coarse, hipStreamSynchronize(st));
  n->count = r1;
  if (added) *added = r1 - r0;
  return FALL3D_ADS_OK;
}

int fall3d_ads_coarse_grain(fall3d_ads_ctx *fine, fall3d_ads_ctx *coarse) {
  if (!fine || !coarse) return FALL3D_ADS_EINVAL;
  if (fine->nx != 2 * coarse->nx || fine->ny != 2 * coarse->ny || fine->nz != coarse->nz || fine->nb != coarse->nb)
    return fail(coarse, FALL3D_ADS_EINVAL, "fine grid must be the 2x horizontal refinement of the coarse grid");
  if (is_fp64(fine) || is_fp64(coarse))
    return fail(coarse, FALL3D_ADS_EINVAL, "coarse graining runs on FP32 states (read an FP64 state with get_state)");
  HIPCHK(fine, hipStreamSynchronize(fine->stream));
  const size_t nc = (size_t)coarse->nx * coarse->ny * coarse->nz * coarse->nb;
  hipLaunchKernelGGL(k_coarse_grain, dim3((unsigned)std::min<size_t>((nc + 255) / 256, 65536)), dim3(256), 0,
                     coarse->stream, fine->chi, coarse->chi, coarse->nx, coarse->ny, coarse->nz, coarse->nb);
  HIPCHK(coarse, hipGetLastError());
  HIPCHK(coarse, hipStreamSynchronize(coarse->stream));
  return FALL3D_ADS_OK;
}

}  // extern "C"
