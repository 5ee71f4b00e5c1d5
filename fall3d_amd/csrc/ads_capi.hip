// ads_capi.hip -- the extern "C" ABI of libfall3d_ads_hip.so (include/fall3d_ads.h).
//
// Host-side orchestration of the gfx950 kernels: context and device memory,
// host<->device staging for the drop-in MODULE ADS entry points, the
// resident F3D_time_step, reductions, CFL, budget and per-kernel timing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "ads_context.h"
#include "ads_fp64.h"

using namespace f3d;

namespace f3d {
namespace host {

int fail(fall3d_ads_ctx *c, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return code;
}

int ensure_stage(fall3d_ads_ctx *c, size_t bytes) {
  if (c->dstage_bytes >= bytes) return FALL3D_ADS_OK;
  dfree(c->dstage);
  c->dstage = nullptr;
  c->dstage_bytes = 0;
  HIPCHK(c, hipMalloc(&c->dstage, bytes));
  c->dstage_bytes = bytes;
  return FALL3D_ADS_OK;
}

// ---- profiling -------------------------------------------------------------
hipEvent_t get_event(fall3d_ads_ctx *c) {
  if (!c->evpool.empty()) {
    hipEvent_t e = c->evpool.back();
    c->evpool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

struct ProfScope {
  fall3d_ads_ctx *c;
  int kid;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(fall3d_ads_ctx *ctx, int k) : c(ctx), kid(k) {
    if (c->prof) {
      a = get_event(c);
      b = get_event(c);
      if (a) (void)hipEventRecord(a, c->stream);
    }
  }
  ~ProfScope() {
    if (c->prof && a && b) {
      (void)hipEventRecord(b, c->stream);
      c->pending.push_back({kid, a, b});
    }
  }
};

void drain_events(fall3d_ads_ctx *c) {
  if (c->pending.empty()) return;
  (void)hipStreamSynchronize(c->stream);
  for (auto &p : c->pending) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      c->kt_ms[p.kid] += ms;
      c->kt_n[p.kid] += 1;
    }
    c->evpool.push_back(p.a);
    c->evpool.push_back(p.b);
  }
  c->pending.clear();
}

// ---- tiling -------------------------------------------------------------------
const int kCpt[] = {2, 3, 4, 6, 8, 12, 16, 20};

Tiling choose_tiling(const fall3d_ads_ctx *c, int dir) {
  Tiling t;
  const int n = axis_n(c, dir);
  if (dir == 0) {
    // C cells per lane, S = n/C + 1 segments per line. Up to 64 segments:
    // one wave per line (several short lines per wave); up to 256: wpl waves
    // of one workgroup per line, trading their edge cells through LDS. All
    // bins of a line in the same wave(s) unless the grid is too small.
    // Default C: the smallest with one wave per line if that is at most 8
    // cells per lane (register pressure grows with C), else 6.
    int C = 6;
    if (c->user_cpt[0]) {
      C = c->user_cpt[0];
    } else {
      for (int cc : kCpt) {
        if (n / cc + 1 <= 64) {
          if (cc <= 8) C = cc;
          break;
        }
      }
    }
    t.cpt = C;
    t.S = n / C + 1;  // the last segment holds face n+1 (possibly no cell)
    t.wpl = (t.S + 63) / 64;
    if (t.wpl > 4) {  // more than 256 lanes per line: unsupported, reported by the caller
      t.S = 0;
      return t;
    }
    int lpb;  // lines per workgroup
    if (t.wpl <= 1) {
      t.wpl = 1;
      t.block = 256;
      lpb = 4 * (t.S ? 64 / t.S : 1);
    } else {
      lpb = 4 / t.wpl;
      t.block = 64 * t.wpl * lpb;
    }
    t.TX = lpb;
    const long nlines = (long)c->ny * c->nz;
    t.gx = (int)((nlines + t.TX - 1) / t.TX);
    int groups = 1;
    if (c->user_bpg[0] > 0) {
      groups = (c->nb + c->user_bpg[0] - 1) / c->user_bpg[0];
    } else {
      const long waves = (long)t.gx * (t.block / 64);
      const long target = 4096;
      if (waves < target) groups = (int)std::min<long>(c->nb, (target + waves - 1) / std::max<long>(waves, 1));
    }
    t.bpg = (c->nb + groups - 1) / groups;
    t.groups = (c->nb + t.bpg - 1) / t.bpg;
    t.ntrans = 0;
  } else {
    // streaming kernel: lanes = 64 x-columns, waves = bins of those columns
    int bpg = c->user_bpg[dir] > 0 ? std::min(c->user_bpg[dir], 8) : 8;
    int groups = (c->nb + bpg - 1) / bpg;
    bpg = (c->nb + groups - 1) / groups;
    t.cpt = c->user_cpt[dir] ? c->user_cpt[dir] : ((c->tm == 1) ? 2 : 1);  // bins per thread
    t.S = 1;
    t.TX = 64;
    t.bpg = bpg;
    t.groups = groups;
    t.block = 64 * ((bpg + t.cpt - 1) / t.cpt);
    t.ntrans = (c->nx + 63) / 64;
    t.gx = groups * t.ntrans * (dir == 1 ? c->nz : c->ny);
  }
  (void)n;
  return t;
}

int refresh_tiling(fall3d_ads_ctx *c) {
  c->version++;
  if (is_fp64(c)) {
    // the FP64 sweeps: one thread per line, a capped grid of 256-thread workgroups
    for (int d = 0; d < 3; ++d) {
      const size_t need = (size_t)sweep64_blocks(axis_lines(c, d), c->nb);
      if (c->part_cap[d + 1] < need) {
        dfree(c->part[d + 1]);
        c->part[d + 1] = nullptr;
        c->part_cap[d + 1] = 0;
        HIPCHK(c, dalloc(&c->part[d + 1], need * 4));
        c->part_cap[d + 1] = need;
      }
    }
    return FALL3D_ADS_OK;
  }
  for (int d = 0; d < 3; ++d) {
    Tiling t = choose_tiling(c, d);
    if (d == 0 && (t.S <= 0 || t.cpt < 2 || !sweep_cpt_supported(t.cpt)))
      return fail(c, FALL3D_ADS_EINVAL, "no supported x tiling (nx=%d cpt=%d; nx < 256 cpt is required)",
                  axis_n(c, d), t.cpt);
    c->tl[d] = t;
    const size_t need = (d == 0) ? (size_t)t.gx * t.groups : (size_t)t.gx;
    if (c->part_cap[d + 1] < need) {
      dfree(c->part[d + 1]);
      dfree(c->rpart[d + 1]);
      c->part[d + 1] = c->rpart[d + 1] = nullptr;
      c->part_cap[d + 1] = 0;
      HIPCHK(c, dalloc(&c->part[d + 1], need * 4));
      HIPCHK(c, dalloc(&c->rpart[d + 1], need * 4));
      c->part_cap[d + 1] = need;
    }
  }
  return FALL3D_ADS_OK;
}

// ---- sweeps ---------------------------------------------------------------------
struct SweepOut {
  int nblocks = 0;
};

// dt/dxb tables of the three axes for this dt, refreshed (on the stream)
// only when dt or the grid changes.
int ensure_dt(fall3d_ads_ctx *c, double dt) {
  if (c->dtib_ok && c->dtib_dt == dt) return FALL3D_ADS_OK;
  for (int d = 0; d < 3; ++d) {
    hipError_t e = dt_over_dxb(c->dxb64[d], axis_n(c, d) + 3, dt, c->tm, c->dttab[d], c->dttab64[d], c->stream);
    if (e == hipSuccess)
      e = iter_table(c->dttab[d], c->hdxp[d], c->dxb64[d], axis_n(c, d), c->tm == 1 ? 1 : 4, c->itab[d], c->stream);
    if (e != hipSuccess) return fail(c, FALL3D_ADS_EHIP, "dt/dxb: %s", hipGetErrorString(e));
  }
  c->dtib_dt = dt;
  c->dtib_ok = true;
  return FALL3D_ADS_OK;
}

void make_sweep_args(fall3d_ads_ctx *c, int dir, double dt, double stime, int write_ghosts, const float *vel1,
                     const float *vel2, const float *kd, SweepArgs &a) {
  const Tiling &t = c->tl[dir];
  memset(&a, 0, sizeof(a));
  a.chi = c->chi;
  a.clo = c->comp ? c->clo : nullptr;
  a.nx = c->nx;
  a.ny = c->ny;
  a.nz = c->nz;
  a.n = axis_n(c, dir);
  a.S = t.S;
  a.wpl = t.wpl;
  a.TX = t.TX;
  a.ntrans = t.ntrans;
  a.ntrans2 = t.groups;
  a.gx = t.gx;
  a.bin0 = 0;
  a.nbin = t.bpg;
  a.nbins = c->nb;
  a.tm = c->tm;
  a.bc = c->bc;
  a.write_ghosts = write_ghosts;
  a.dt = (float)dt;
  a.stime = (float)stime;
  a.vel1 = vel1;
  a.vel2 = vel2;
  a.kd = kd;
  a.vs = c->vs;
  a.hdxp = c->hdxp[dir];
  a.idxb = c->idxb[dir];
  {
    const int n3 = axis_n(c, dir) + 3;
    a.dtib = c->dttab[dir];
    a.hdtib = c->dttab[dir] + n3;
    a.ftib = c->dttab[dir] + 2 * n3;
    a.ftib64 = c->dttab64[dir];
    a.itab = c->itab[dir];
  }
  a.dxb64 = c->dxb64[dir];
  if (dir == 0) {
    a.areaA = c->dxb64[1];
    a.areaB = c->dxb64[2];
  } else if (dir == 1) {
    a.areaA = c->dxb64[0];
    a.areaB = c->dxb64[2];
  } else {
    a.areaA = c->dxb64[0];
    a.areaB = c->dxb64[1];
  }
  a.acum = c->acum;
  a.partials = c->part[dir + 1];
  a.rpartials = c->rpart[dir + 1];
  if (dir < 2 && (c->fout_on || (c->learned && c->learned_flux_limit > 0.0))) {
    // face fluxes: accumulated for training rows, overwritten for the
    // learned step's flux-relative clamp
    a.fout = c->fout[dir];
    a.fscale = (float)(c->tm == 1 ? dt : dt / 6.0);
    a.fout_add = c->fout_on ? 1 : 0;
  }
}

void make_sweep_args(fall3d_ads_ctx *c, int dir, double dt, double stime, int write_ghosts, SweepArgs &a) {
  const float *v1 = dir == 0 ? c->u1 : (dir == 1 ? c->v1 : c->w1);
  const float *v2 = dir == 0 ? c->u2 : (dir == 1 ? c->v2 : c->w2);
  const float *kd = dir == 0 ? c->k1 : (dir == 1 ? c->k2 : c->k3);
  make_sweep_args(c, dir, dt, stime, write_ghosts, v1, v2, kd, a);
}

// partials, rpartials: per-workgroup partial sums of the launch and of its
// rounding residual (nullptr: the context's own buffers of that sweep)
int run_sweep(fall3d_ads_ctx *c, int dir, double dt, double stime, const float *vel1, const float *vel2,
              const float *kd, int write_ghosts, SweepOut *out, double *partials = nullptr,
              double *rpartials = nullptr) {
  {
    const int rc = ensure_dt(c, dt);
    if (rc) return rc;
  }
  const Tiling &t = c->tl[dir];
  SweepArgs a;
  make_sweep_args(c, dir, dt, stime, write_ghosts, vel1, vel2, kd, a);
  if (partials) a.partials = partials;
  if (rpartials) a.rpartials = rpartials;
  dim3 grid = (dir == 0) ? dim3(t.gx, t.groups) : dim3(t.gx);
  dim3 block(t.block);
  {
    ProfScope ps(c, KID_SWEEP_X + dir);
    hipError_t e = launch_sweep(dir, t.cpt, c->limiter, c->comp != 0, c->tm == 1 ? 1 : 4, a, grid, block,
                                c->stream);
    if (e != hipSuccess) return fail(c, FALL3D_ADS_EHIP, "sweep %d launch: %s", dir, hipGetErrorString(e));
  }
  out->nblocks = (dir == 0) ? t.gx * t.groups : t.gx;
  return FALL3D_ADS_OK;
}

// dst = src with its two fastest axes swapped (A x B planes, K of them),
// in launches of at most 65535 planes
hipError_t transpose64(const double *src, double *dst, int A, int B, size_t K, hipStream_t st) {
  const size_t plane = (size_t)A * B;
  for (size_t k0 = 0; k0 < K; k0 += 65535) {
    const int kk = (int)(K - k0 < 65535 ? K - k0 : 65535);
    hipError_t e = transpose_to_f64(src + k0 * plane, 8, dst + k0 * plane, A, B, kk, st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// FP64 state: one sweep with the given FP64 winds (vel1, vel2: lerped by
// stime), diffusion and, for z, settling velocity. The x sweep runs on
// (x, y)-transposed copies of the state, winds and diffusion, and the state
// is transposed back after it.
int run_sweep64(fall3d_ads_ctx *c, int dir, double dt, double stime, const double *vel1, const double *vel2,
                const double *kd, SweepOut *out) {
  Sweep64Args a;
  memset(&a, 0, sizeof(a));
  a.c = c->c64;
  a.nx = c->nx;
  a.ny = c->ny;
  a.nz = c->nz;
  a.nb = c->nb;
  a.n = axis_n(c, dir);
  a.tm = c->tm;
  a.bc = c->bc;
  a.nlines = axis_lines(c, dir);
  a.dt = dt;
  a.stime = stime;
  a.vel1 = vel1;
  a.vel2 = vel2;
  a.kd = kd;
  a.vs = c->vsd;
  a.dxp = c->dxp64[dir];
  a.dxb = c->dxb64[dir];
  a.areaA = c->dxb64[dir == 0 ? 1 : 0];
  a.areaB = c->dxb64[dir == 2 ? 1 : 2];
  a.acum = c->acum;
  a.partials = c->part[dir + 1];
  const int nblocks = sweep64_blocks(a.nlines, c->nb);
  {
    ProfScope ps(c, KID_SWEEP_X + dir);
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const size_t planes = (size_t)(nz + 4) * c->nb;
    if (dir == 0) {
      hipError_t e = transpose64(c->c64, c->c64t, nx + 4, ny + 4, planes, c->stream);
      if (e == hipSuccess) e = transpose64(vel1, c->u1t, nx + 3, ny, nz, c->stream);
      if (e == hipSuccess && vel2 != vel1) e = transpose64(vel2, c->u2t, nx + 3, ny, nz, c->stream);
      if (e == hipSuccess) e = transpose64(kd, c->k1t, nx + 4, ny, nz, c->stream);
      if (e != hipSuccess) return fail(c, FALL3D_ADS_EHIP, "FP64 x-sweep transposes: %s", hipGetErrorString(e));
      a.c = c->c64t;
      a.vel1 = c->u1t;
      a.vel2 = vel2 != vel1 ? c->u2t : c->u1t;
      a.kd = c->k1t;
    }
    hipError_t e = launch_sweep64(dir, c->limiter, a, nblocks, c->stream);
    if (e != hipSuccess) return fail(c, FALL3D_ADS_EHIP, "FP64 sweep %d launch: %s", dir, hipGetErrorString(e));
    if (dir == 0) {
      e = transpose64(c->c64t, c->c64, ny + 4, nx + 4, planes, c->stream);
      if (e != hipSuccess) return fail(c, FALL3D_ADS_EHIP, "FP64 x-sweep transpose: %s", hipGetErrorString(e));
    }
  }
  out->nblocks = nblocks;
  return FALL3D_ADS_OK;
}

void add_sweep_jobs(ReduceJob &j, const fall3d_ads_ctx *c, int dir, int nblocks, const double *partials = nullptr,
                    const double *rpartials = nullptr) {
  // partial slots: 0 ref beg, 1 ref end, 2 ledger beg, 3 ledger end
  const int dst[4] = {ACC_FLUX + 2 * dir, ACC_FLUX + 2 * dir + 1, ACC_LEDGER + 2 * dir, ACC_LEDGER + 2 * dir + 1};
  j.part[j.count] = partials ? partials : c->part[dir + 1];
  j.nblocks[j.count] = nblocks;
  j.nslot[j.count] = 4;
  for (int s = 0; s < 4; ++s) j.dst[j.count][s] = dst[s];
  j.count++;
  if (is_fp64(c)) return;  // no rounding residual to account for
  // rounding residual of the FP32 state's final updates (zero for the compensated state)
  j.part[j.count] = rpartials ? rpartials : c->rpart[dir + 1];
  j.nblocks[j.count] = nblocks;
  j.nslot[j.count] = 1;
  j.dst[j.count][0] = ACC_RESIDUAL + dir;
  j.count++;
}

int run_finalize(fall3d_ads_ctx *c, const ReduceJob &j) {
  ProfScope ps(c, KID_REDUCE);
  hipError_t e = finalize(c->acc, j, c->stream);
  if (e != hipSuccess) return fail(c, FALL3D_ADS_EHIP, "finalize: %s", hipGetErrorString(e));
  return FALL3D_ADS_OK;
}

// upload a host array of the context kind into device staging
int upload(fall3d_ads_ctx *c, const void *h, size_t nelem, void **dptr) {
  const size_t bytes = nelem * (size_t)c->kind;
  int rc = ensure_stage(c, bytes);
  if (rc) return rc;
  HIPCHK(c, hipMemcpyAsync(c->dstage, h, bytes, hipMemcpyHostToDevice, c->stream));
  *dptr = c->dstage;
  return FALL3D_ADS_OK;
}

int need_grid(fall3d_ads_ctx *c) {
  if (!c->grid_ok) return fail(c, FALL3D_ADS_ESTATE, "fall3d_ads_set_grid has not been called");
  return FALL3D_ADS_OK;
}

size_t n_u(const fall3d_ads_ctx *c) { return (size_t)(c->nx + 3) * c->ny * c->nz; }
size_t n_v(const fall3d_ads_ctx *c) { return (size_t)(c->ny + 3) * c->nx * c->nz; }
size_t n_w(const fall3d_ads_ctx *c) { return (size_t)(c->nz + 3) * c->nx * c->ny; }
size_t n_k1(const fall3d_ads_ctx *c) { return (size_t)(c->nx + 4) * c->ny * c->nz; }
size_t n_k2(const fall3d_ads_ctx *c) { return (size_t)(c->ny + 4) * c->nx * c->nz; }
size_t n_k3(const fall3d_ads_ctx *c) { return (size_t)(c->nz + 4) * c->nx * c->ny; }
size_t n_vs(const fall3d_ads_ctx *c) { return (size_t)c->nx * c->ny * (c->nz + 3) * c->nb; }

// reference layout -> device layout for each coefficient field
// Field uploads: host kind -> FP32 device layout d and, for the FP64 state,
// the same layout in FP64 (d64). Fields kept in the reference layout copy;
// the others swap their two fastest axes (x first).
int up_field(fall3d_ads_ctx *c, const void *h, size_t n, int A, int B, int K, float *d, double *d64) {
  void *s;
  int rc = upload(c, h, n, &s);
  if (rc) return rc;
  if (A == 0) {
    HIPCHK(c, to_f32(s, c->kind, d, n, c->stream));
    if (d64) HIPCHK(c, to_f64(s, c->kind, d64, n, c->stream));
  } else {
    HIPCHK(c, transpose_to_f32(s, c->kind, d, A, B, K, c->stream));
    if (d64) HIPCHK(c, transpose_to_f64(s, c->kind, d64, A, B, K, c->stream));
  }
  return FALL3D_ADS_OK;
}
int up_u(fall3d_ads_ctx *c, const void *h, float *d, double *d64) { return up_field(c, h, n_u(c), 0, 0, 0, d, d64); }
int up_v(fall3d_ads_ctx *c, const void *h, float *d, double *d64) {
  return up_field(c, h, n_v(c), c->ny + 3, c->nx, c->nz, d, d64);
}
int up_w(fall3d_ads_ctx *c, const void *h, float *d, double *d64) {
  return up_field(c, h, n_w(c), c->nz + 3, c->nx * c->ny, 1, d, d64);
}
int up_k1(fall3d_ads_ctx *c, const void *h) { return up_field(c, h, n_k1(c), 0, 0, 0, c->k1, c->k1d); }
int up_k2(fall3d_ads_ctx *c, const void *h) {
  return up_field(c, h, n_k2(c), c->ny + 4, c->nx, c->nz, c->k2, c->k2d);
}
int up_k3(fall3d_ads_ctx *c, const void *h) {
  return up_field(c, h, n_k3(c), c->nz + 4, c->nx * c->ny, 1, c->k3, c->k3d);
}
int up_vs(fall3d_ads_ctx *c, const void *h) { return up_field(c, h, n_vs(c), 0, 0, 0, c->vs, c->vsd); }

int up_state(fall3d_ads_ctx *c, const void *h) {
  void *s;
  int rc = upload(c, h, c->ncell_c, &s);
  if (rc) return rc;
  if (is_fp64(c))
    HIPCHK(c, to_f64(s, c->kind, c->c64, c->ncell_c, c->stream));
  else
    HIPCHK(c, state_in(s, c->kind, c->chi, c->comp ? c->clo : nullptr, c->nx + 4, state_pitch(c->nx),
                       (size_t)(c->ny + 4) * (c->nz + 4) * c->nb, c->stream));
  return FALL3D_ADS_OK;
}

int down_state(fall3d_ads_ctx *c, void *h) {
  const size_t bytes = c->ncell_c * (size_t)c->kind;
  int rc = ensure_stage(c, bytes);
  if (rc) return rc;
  if (is_fp64(c))
    HIPCHK(c, from_f64(c->c64, c->dstage, c->kind, c->ncell_c, c->stream));
  else
    HIPCHK(c, state_out(c->chi, c->comp ? c->clo : nullptr, c->dstage, c->kind, c->nx + 4, state_pitch(c->nx),
                        (size_t)(c->ny + 4) * (c->nz + 4) * c->nb, c->stream));
  HIPCHK(c, hipMemcpyAsync(h, c->dstage, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FALL3D_ADS_OK;
}

int up_acum(fall3d_ads_ctx *c, const void *h) {
  void *s;
  int rc = upload(c, h, c->nacum, &s);
  if (rc) return rc;
  HIPCHK(c, to_f64(s, c->kind, c->acum, c->nacum, c->stream));
  return FALL3D_ADS_OK;
}

int down_acum(fall3d_ads_ctx *c, void *h) {
  if (c->kind == 8) {
    HIPCHK(c, hipMemcpyAsync(h, c->acum, c->nacum * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return FALL3D_ADS_OK;
  }
  std::vector<double> tmp(c->nacum);
  HIPCHK(c, hipMemcpyAsync(tmp.data(), c->acum, c->nacum * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  float *o = (float *)h;
  for (size_t i = 0; i < c->nacum; ++i) o[i] = (float)tmp[i];
  return FALL3D_ADS_OK;
}

double host_val(const void *p, int kind, size_t i) {
  return kind == 8 ? ((const double *)p)[i] : (double)((const float *)p)[i];
}

int set_acc(fall3d_ads_ctx *c, int idx, const double *vals, int n) {
  HIPCHK(c, hipMemcpyAsync(c->acc + idx, vals, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  return FALL3D_ADS_OK;
}

int get_acc(fall3d_ads_ctx *c, int idx, double *vals, int n) {
  HIPCHK(c, hipMemcpyAsync(vals, c->acc + idx, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FALL3D_ADS_OK;
}

// one drop-in sweep: flux pair in/out through acc[ACC_FLUX + 2 dir ..]
int dropin_sweep(fall3d_ads_ctx *c, int dir, double *fbeg, double *fend, double dt, void *cst, const void *kd,
                 const void *w1, const void *w2, double stime, const void *vsh, void *acum) {
  int rc;
  if ((rc = need_grid(c))) return rc;
  if (!cst || !kd || !fbeg || !fend) return fail(c, FALL3D_ADS_EINVAL, "null argument");
  if ((rc = up_state(c, cst))) return rc;
  const float *vel1 = nullptr, *vel2 = nullptr, *kdd = nullptr;
  const double *vd1 = nullptr, *vd2 = nullptr, *kd64 = nullptr;
  if (dir == 0) {
    if ((rc = up_u(c, c->hu, c->u1, c->u1d))) return rc;
    if ((rc = up_k1(c, kd))) return rc;
    vel1 = vel2 = c->u1;
    vd1 = vd2 = c->u1d;
    kdd = c->k1;
    kd64 = c->k1d;
    stime = 0.0;
  } else if (dir == 1) {
    if ((rc = up_v(c, c->hv, c->v1, c->v1d))) return rc;
    if ((rc = up_k2(c, kd))) return rc;
    vel1 = vel2 = c->v1;
    vd1 = vd2 = c->v1d;
    kdd = c->k2;
    kd64 = c->k2d;
    stime = 0.0;
  } else {
    if ((rc = up_w(c, w1, c->w1, c->w1d))) return rc;
    if ((rc = up_w(c, w2, c->w2, c->w2d))) return rc;
    if ((rc = up_k3(c, kd))) return rc;
    if ((rc = up_vs(c, vsh))) return rc;
    if ((rc = up_acum(c, acum))) return rc;
    vel1 = c->w1;
    vel2 = c->w2;
    vd1 = c->w1d;
    vd2 = c->w2d;
    kdd = c->k3;
    kd64 = c->k3d;
  }
  double pair[2] = {*fbeg, *fend};
  if ((rc = set_acc(c, ACC_FLUX + 2 * dir, pair, 2))) return rc;
  SweepOut so;
  if (is_fp64(c)) {
    if ((rc = ensure_dt(c, dt))) return rc;
    if ((rc = run_sweep64(c, dir, dt, stime, vd1, vd2, kd64, &so))) return rc;
  } else if ((rc = run_sweep(c, dir, dt, stime, vel1, vel2, kdd, 1, &so))) {
    return rc;
  }
  ReduceJob j;
  memset(&j, 0, sizeof(j));
  add_sweep_jobs(j, c, dir, so.nblocks);
  if ((rc = run_finalize(c, j))) return rc;
  if ((rc = down_state(c, cst))) return rc;
  if (dir == 2 && (rc = down_acum(c, acum))) return rc;
  if ((rc = get_acc(c, ACC_FLUX + 2 * dir, pair, 2))) return rc;
  *fbeg = pair[0];
  *fend = pair[1];
  // the meteo buffers now hold per-call data
  c->meteo_ok = false;
  return FALL3D_ADS_OK;
}

// ---- allocation-free step ---------------------------------------------------
// Per-workgroup partial sums of fall3d_ads_step_noalloc live in this array
// (one copy per GPU, held until the process ends). Host code never uses the
// symbol itself: its device address is looked up once per device with
// hipGetSymbolAddress and kept in noalloc_ptr.
constexpr size_t kNoAllocScratchDoubles = (size_t)1 << 21;  // 16 MiB
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex noalloc_mu;
static double *noalloc_ptr[kMaxDevices] = {};

// device address of this device's scratch array (the current device)
int noalloc_scratch(fall3d_ads_ctx *c, double **out) {
  int dev = -1;
  HIPCHK(c, hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(c, FALL3D_ADS_EINVAL, "device %d out of range", dev);
  std::lock_guard<std::mutex> lk(noalloc_mu);
  if (!noalloc_ptr[dev]) {
    double *p = nullptr;
    hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(c, FALL3D_ADS_EHIP, "scratch symbol address: %s", hipGetErrorString(e));
    noalloc_ptr[dev] = p;
  }
  *out = noalloc_ptr[dev];
  return FALL3D_ADS_OK;
}

// One F3D_time_step. part (nullptr: the context's buffers) gives the
// per-workgroup partial-sum buffers of the source (0), the x, y, z sweeps
// (1..3) and the sweeps' rounding residuals (4..6).
int step_impl(fall3d_ads_ctx *c, double dt, double stime, int iiter, double *const *part) {
  int rc;
  if ((rc = need_grid(c))) return rc;
  if (!c->meteo_ok) return fail(c, FALL3D_ADS_ESTATE, "fall3d_ads_set_meteo has not been called");
  if (c->learned && (rc = learned_pre_step(c, dt, stime))) return rc;
  ReduceJob j;
  memset(&j, 0, sizeof(j));
  if (c->nsrc > 0) {
    int nblk = 0;
    double *sp = part ? part[0] : c->src_part;
    {
      ProfScope ps(c, KID_SOURCE);
      hipError_t e = is_fp64(c) ? source_inject64(c->c64, c->src_idx, c->src_s, c->src_vol, c->src_meas, c->nsrc,
                                                  dt, sp, &nblk, c->stream)
                                : source_inject(c->chi, c->comp ? c->clo : nullptr, c->src_idx, c->src_rate,
                                                c->src_meas, c->nsrc, dt, sp, &nblk, c->stream);
      if (e != hipSuccess) return fail(c, FALL3D_ADS_EHIP, "source: %s", hipGetErrorString(e));
    }
    j.part[j.count] = sp;
    j.nblocks[j.count] = nblk;
    j.nslot[j.count] = 1;
    j.dst[j.count][0] = ACC_INJECTED;
    j.count++;
  }
  SweepOut so;
  const int order[3] = {(iiter % 2 == 0) ? 0 : 1, (iiter % 2 == 0) ? 1 : 0, 2};
  if (is_fp64(c)) {
    if ((rc = ensure_dt(c, dt))) return rc;
    for (int q = 0; q < 3; ++q) {
      const int d = order[q];
      const double *v1 = d == 0 ? c->u1d : (d == 1 ? c->v1d : c->w1d);
      const double *v2 = d == 0 ? c->u2d : (d == 1 ? c->v2d : c->w2d);
      const double *kd = d == 0 ? c->k1d : (d == 1 ? c->k2d : c->k3d);
      if ((rc = run_sweep64(c, d, dt, stime, v1, v2, kd, &so))) return rc;
      add_sweep_jobs(j, c, d, so.nblocks);
    }
    return run_finalize(c, j);
  }
  for (int q = 0; q < 3; ++q) {
    const int d = order[q];
    const float *v1 = d == 0 ? c->u1 : (d == 1 ? c->v1 : c->w1);
    const float *v2 = d == 0 ? c->u2 : (d == 1 ? c->v2 : c->w2);
    const float *kd = d == 0 ? c->k1 : (d == 1 ? c->k2 : c->k3);
    double *pp = part ? part[d + 1] : nullptr;
    double *rp = part ? part[d + 4] : nullptr;
    if ((rc = run_sweep(c, d, dt, stime, v1, v2, kd, 1, &so, pp, rp))) return rc;
    add_sweep_jobs(j, c, d, so.nblocks, pp, rp);
  }
  if (c->learned && (rc = learned_post_step(c))) return rc;
  return run_finalize(c, j);
}

}  // namespace host
}  // namespace f3d

using namespace f3d::host;

// =============================================================================
// C ABI
// =============================================================================
extern "C" {

int fall3d_ads_initialize(fall3d_ads_ctx **pctx, int nx, int ny, int nz, int nbins, int limiter, int time_marching,
                          int real_kind, int state_mode) {
  if (!pctx) return FALL3D_ADS_EINVAL;
  *pctx = nullptr;
  fall3d_ads_ctx *c = new (std::nothrow) fall3d_ads_ctx();
  if (!c) return FALL3D_ADS_EINVAL;
  *pctx = c;
  if (nx < 2 || ny < 2 || nz < 2 || nbins < 1)
    return fail(c, FALL3D_ADS_EINVAL, "invalid grid %d x %d x %d x %d (each axis needs >= 2 cells)", nx, ny, nz,
                nbins);
  if (limiter < 1 || limiter > 3) return fail(c, FALL3D_ADS_EINVAL, "invalid limiter %d", limiter);
  if (time_marching != 1 && time_marching != 2)
    return fail(c, FALL3D_ADS_EINVAL, "invalid time_marching %d", time_marching);
  if (real_kind != 4 && real_kind != 8) return fail(c, FALL3D_ADS_EINVAL, "invalid real kind %d", real_kind);
  if (state_mode < 0 || state_mode > 2) return fail(c, FALL3D_ADS_EINVAL, "invalid state mode %d", state_mode);
  {
    // The sweeps address each bin's concentration, and each wind, diffusion
    // and settling-velocity array, through a buffer descriptor whose record
    // count is a 32-bit byte size and drop masked stores with an offset of
    // 2^31: every such array must stay below 2^31 bytes.
    const size_t lim = (size_t)1 << 31;
    const size_t per_bin_c = (size_t)state_pitch(nx) * (ny + 4) * (nz + 4) * sizeof(float);
    const size_t faces = std::max({(size_t)(nx + 4) * ny * nz, (size_t)nx * (ny + 4) * nz, (size_t)nx * ny * (nz + 4)}) *
                         sizeof(float);
    if (per_bin_c >= lim || faces >= lim)
      return fail(c, FALL3D_ADS_EINVAL,
                  "grid %d x %d x %d too large: one bin of the concentration (with ghosts) or one wind/diffusion "
                  "field must stay below 2 GiB",
                  nx, ny, nz);
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(c, FALL3D_ADS_EHIP, "no HIP device available");
  {
    // device address of the allocation-free step's scratch array on this device
    double *scr = nullptr;
    const int rc = noalloc_scratch(c, &scr);
    if (rc) return rc;
  }
  c->nx = nx;
  c->ny = ny;
  c->nz = nz;
  c->nb = nbins;
  c->limiter = limiter;
  c->tm = time_marching;
  c->kind = real_kind;
  c->comp = state_mode;
  HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->own_stream = true;
  c->ncell_c = (size_t)(nx + 4) * (ny + 4) * (nz + 4) * nbins;
  c->ncell_s = (size_t)state_pitch(nx) * (ny + 4) * (nz + 4) * nbins;
  c->nacum = (size_t)(nx + 4) * (ny + 4) * nbins;
  if (is_fp64(c)) {
    HIPCHK(c, dalloc(&c->c64, c->ncell_c));
    HIPCHK(c, hipMemsetAsync(c->c64, 0, c->ncell_c * sizeof(double), c->stream));
    double **m64[10] = {&c->u1d, &c->u2d, &c->v1d, &c->v2d, &c->w1d, &c->w2d, &c->k1d, &c->k2d, &c->k3d, &c->vsd};
    const size_t n64[10] = {n_u(c), n_u(c), n_v(c), n_v(c), n_w(c), n_w(c), n_k1(c), n_k2(c), n_k3(c), n_vs(c)};
    for (int q = 0; q < 10; ++q) HIPCHK(c, dalloc(m64[q], n64[q]));
    HIPCHK(c, dalloc(&c->c64t, c->ncell_c));
    HIPCHK(c, dalloc(&c->u1t, n_u(c)));
    HIPCHK(c, dalloc(&c->u2t, n_u(c)));
    HIPCHK(c, dalloc(&c->k1t, n_k1(c)));
  } else {
    HIPCHK(c, dalloc(&c->chi_base, c->ncell_s + kStateLead + 2));
    c->chi = c->chi_base + kStateLead;
    HIPCHK(c, hipMemsetAsync(c->chi_base, 0, (c->ncell_s + kStateLead + 2) * sizeof(float), c->stream));
  }
  if (c->comp == FALL3D_ADS_STATE_FP32X2) {
    HIPCHK(c, dalloc(&c->clo_base, c->ncell_s + kStateLead + 2));
    c->clo = c->clo_base + kStateLead;
    HIPCHK(c, hipMemsetAsync(c->clo_base, 0, (c->ncell_s + kStateLead + 2) * sizeof(float), c->stream));
  }
  HIPCHK(c, dalloc(&c->acum, c->nacum));
  HIPCHK(c, hipMemsetAsync(c->acum, 0, c->nacum * sizeof(double), c->stream));
  HIPCHK(c, dalloc(&c->u1, n_u(c)));
  HIPCHK(c, dalloc(&c->u2, n_u(c)));
  HIPCHK(c, dalloc(&c->v1, n_v(c)));
  HIPCHK(c, dalloc(&c->v2, n_v(c)));
  HIPCHK(c, dalloc(&c->w1, n_w(c)));
  HIPCHK(c, dalloc(&c->w2, n_w(c)));
  HIPCHK(c, dalloc(&c->k1, n_k1(c)));
  HIPCHK(c, dalloc(&c->k2, n_k2(c)));
  HIPCHK(c, dalloc(&c->k3, n_k3(c)));
  HIPCHK(c, dalloc(&c->vs, n_vs(c)));
  HIPCHK(c, dalloc(&c->acc, ACC_LEN));
  HIPCHK(c, hipMemsetAsync(c->acc, 0, ACC_LEN * sizeof(double), c->stream));
  c->red_cap = 4096;
  HIPCHK(c, dalloc(&c->red_part, c->red_cap));
  for (int d = 0; d < 3; ++d) {
    const int n = axis_n(c, d);
    HIPCHK(c, dalloc(&c->dxp64[d], n + 4));
    HIPCHK(c, dalloc(&c->dxb64[d], n + 3));
    HIPCHK(c, dalloc(&c->hdxp[d], n + 3));
    HIPCHK(c, dalloc(&c->idxb[d], n + 3));
    HIPCHK(c, dalloc(&c->dttab[d], 3 * (n + 3)));
    HIPCHK(c, dalloc(&c->dttab64[d], n + 3));
    HIPCHK(c, dalloc(&c->itab[d], (size_t)8 * iter_rows(n)));
  }
  const size_t ksz = (size_t)real_kind;
  c->hu = calloc(n_u(c), ksz);
  c->hv = calloc(n_v(c), ksz);
  c->hw = calloc(n_w(c), ksz);
  if (!c->hu || !c->hv || !c->hw) return fail(c, FALL3D_ADS_EINVAL, "host allocation failed");
  int rc = refresh_tiling(c);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FALL3D_ADS_OK;
}

int fall3d_ads_release(fall3d_ads_ctx *c) {
  if (!c) return FALL3D_ADS_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto &p : c->pending) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  for (auto e : c->evpool) (void)hipEventDestroy(e);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  learned_release(c);
  void *ptrs[] = {c->chi_base, c->clo_base, c->acum, c->u1, c->u2, c->v1, c->v2, c->w1, c->w2, c->k1, c->k2, c->k3, c->vs,
                  c->src_idx, c->src_rate, c->src_meas, c->src_part, c->acc, c->part[1], c->part[2],
                  c->part[3], c->rpart[1], c->rpart[2], c->rpart[3], c->red_part,
                  c->dstage, c->c64, c->u1d, c->u2d, c->v1d, c->v2d, c->w1d, c->w2d, c->k1d, c->k2d, c->k3d,
                  c->vsd, c->c64t, c->u1t, c->u2t, c->k1t, c->src_s, c->src_vol};  // part[0] aliases src_part
  for (void *p : ptrs) dfree(p);
  for (int d = 0; d < 3; ++d) {
    dfree(c->dxp64[d]);
    dfree(c->dxb64[d]);
    dfree(c->hdxp[d]);
    dfree(c->idxb[d]);
    dfree(c->dttab[d]);
    dfree(c->dttab64[d]);
    dfree(c->itab[d]);
  }
  free(c->hu);
  free(c->hv);
  free(c->hw);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return FALL3D_ADS_OK;
}

const char *fall3d_ads_last_error(const fall3d_ads_ctx *c) { return c ? c->err.c_str() : "null context"; }

int fall3d_ads_set_grid(fall3d_ads_ctx *c, const void *dX1_p, const void *dX1_b, const void *dX2_p,
                        const void *dX2_b, const void *dX3_p, const void *dX3_b) {
  if (!c) return FALL3D_ADS_EINVAL;
  const void *P[3] = {dX1_p, dX2_p, dX3_p};
  const void *B[3] = {dX1_b, dX2_b, dX3_b};
  for (int d = 0; d < 3; ++d) {
    if (!P[d] || !B[d]) return fail(c, FALL3D_ADS_EINVAL, "null grid array");
    const int n = axis_n(c, d);
    std::vector<double> p(n + 4), b(n + 3);
    for (int i = 0; i < n + 4; ++i) p[i] = host_val(P[d], c->kind, i);
    for (int i = 0; i < n + 3; ++i) b[i] = host_val(B[d], c->kind, i);
    for (int i = 0; i < n + 4; ++i)
      if (!(p[i] > 0.0)) return fail(c, FALL3D_ADS_EINVAL, "non-positive dX%d_p at index %d", d + 1, i);
    for (int i = 0; i < n + 3; ++i)
      if (!(b[i] > 0.0)) return fail(c, FALL3D_ADS_EINVAL, "non-positive dX%d_b at index %d", d + 1, i);
    // 0.5/dxp(f-1) for faces f = 0..n+2 (dxp position f-1 sits at index f)
    std::vector<float> h(n + 3), ib(n + 3);
    for (int f = 0; f < n + 3; ++f) h[f] = (float)(0.5 / p[f]);
    for (int i = 0; i < n + 3; ++i) ib[i] = (float)(1.0 / b[i]);
    HIPCHK(c, hipMemcpy(c->dxp64[d], p.data(), sizeof(double) * (n + 4), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->dxb64[d], b.data(), sizeof(double) * (n + 3), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->hdxp[d], h.data(), sizeof(float) * (n + 3), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->idxb[d], ib.data(), sizeof(float) * (n + 3), hipMemcpyHostToDevice));
    c->hdxp64[d] = p;
    c->hdxb64[d] = b;
  }
  c->grid_ok = true;
  c->dtib_ok = false;
  // the per-axis tables (dt/dxb, iteration factors) and every captured
  // graph or batched-launch table built from the old metrics are stale
  c->version++;
  return FALL3D_ADS_OK;
}

int fall3d_ads_get_vcomps(fall3d_ads_ctx *c, void **u, void **v, void **w) {
  if (!c || !u || !v || !w) return FALL3D_ADS_EINVAL;
  *u = c->hu;
  *v = c->hv;
  *w = c->hw;
  return FALL3D_ADS_OK;
}

int fall3d_ads_update_horizontal_vcomps(fall3d_ads_ctx *c, const void *u1, const void *u2, const void *v1,
                                        const void *v2, double stime) {
  if (!c || !u1 || !u2 || !v1 || !v2) return FALL3D_ADS_EINVAL;
  // interpolate on the GPU in the reference layouts (FP32 results), then
  // refresh the host CB%u / CB%v arrays the caller can read
  int rc;
  const size_t nu = n_u(c), nv = n_v(c);
  float *a = nullptr, *b = nullptr;
  HIPCHK(c, dalloc(&a, std::max(nu, nv)));
  HIPCHK(c, dalloc(&b, std::max(nu, nv)));
  const void *src[2][2] = {{u1, u2}, {v1, v2}};
  void *dst[2] = {c->hu, c->hv};
  const size_t cnt[2] = {nu, nv};
  for (int q = 0; q < 2; ++q) {
    void *s;
    if ((rc = upload(c, src[q][0], cnt[q], &s))) { dfree(a); dfree(b); return rc; }
    HIPCHK(c, to_f32(s, c->kind, a, cnt[q], c->stream));
    if ((rc = upload(c, src[q][1], cnt[q], &s))) { dfree(a); dfree(b); return rc; }
    HIPCHK(c, to_f32(s, c->kind, b, cnt[q], c->stream));
    HIPCHK(c, lerp(a, b, a, (float)stime, cnt[q], c->stream));
    if ((rc = ensure_stage(c, cnt[q] * 8))) { dfree(a); dfree(b); return rc; }
    HIPCHK(c, from_f32(a, c->dstage, c->kind, cnt[q], c->stream));
    HIPCHK(c, hipMemcpyAsync(dst[q], c->dstage, cnt[q] * c->kind, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  dfree(a);
  dfree(b);
  return FALL3D_ADS_OK;
}

int fall3d_ads_solve_along_x(fall3d_ads_ctx *c, double *W_flux, double *E_flux, double dt, void *cst,
                             const void *k1) {
  if (!c) return FALL3D_ADS_EINVAL;
  return dropin_sweep(c, 0, W_flux, E_flux, dt, cst, k1, nullptr, nullptr, 0.0, nullptr, nullptr);
}

int fall3d_ads_solve_along_y(fall3d_ads_ctx *c, double *S_flux, double *N_flux, double dt, void *cst,
                             const void *k2) {
  if (!c) return FALL3D_ADS_EINVAL;
  return dropin_sweep(c, 1, S_flux, N_flux, dt, cst, k2, nullptr, nullptr, 0.0, nullptr, nullptr);
}

int fall3d_ads_solve_along_z(fall3d_ads_ctx *c, double *D_flux, double *U_flux, double dt, void *cst, void *acum,
                             const void *k3, const void *w1, const void *w2, double stime, const void *vs) {
  if (!c) return FALL3D_ADS_EINVAL;
  if (!acum || !w1 || !w2 || !vs) return fail(c, FALL3D_ADS_EINVAL, "null argument");
  return dropin_sweep(c, 2, D_flux, U_flux, dt, cst, k3, w1, w2, stime, vs, acum);
}

int fall3d_ads_solve_1d(fall3d_ads_ctx *c, void *cl, const void *dxp, const void *dxb, const void *u,
                        const void *k, double dt, const double bvalue[2], double flux[2], const int bcode[2], int ips,
                        int ipe, int limiter) {
  if (!c) return FALL3D_ADS_EINVAL;
  if (!cl || !dxp || !dxb || !u || !k || !bvalue || !flux || !bcode) return fail(c, FALL3D_ADS_EINVAL, "null argument");
  const int n = ipe - ips + 1;
  if (n < 2 || n > 1500) return fail(c, FALL3D_ADS_EINVAL, "solve_1d: line length %d out of range [2,1500]", n);
  if (limiter < 1 || limiter > 3) return fail(c, FALL3D_ADS_EINVAL, "invalid limiter %d", limiter);
  for (int s = 0; s < 2; ++s)
    if (bcode[s] < -1 || bcode[s] > 2) return fail(c, FALL3D_ADS_EINVAL, "invalid bcode %d", bcode[s]);
  std::vector<double> hc(n + 4), hp(n + 4), hb(n + 3), hu(n + 3), hk(n + 4);
  for (int i = 0; i < n + 4; ++i) {
    hc[i] = host_val(cl, c->kind, i);
    hp[i] = host_val(dxp, c->kind, i);
    hk[i] = host_val(k, c->kind, i);
  }
  for (int i = 0; i < n + 3; ++i) {
    hb[i] = host_val(dxb, c->kind, i);
    hu[i] = host_val(u, c->kind, i);
  }
  double *d = nullptr;
  const size_t tot = 4 * (size_t)(n + 4) + (n + 3) * 2 + 2;
  HIPCHK(c, dalloc(&d, tot));
  double *dc = d, *dp = dc + (n + 4), *dk = dp + (n + 4), *db = dk + (n + 4), *du = db + (n + 3), *df = du + (n + 3);
  HIPCHK(c, hipMemcpyAsync(dc, hc.data(), 8 * (n + 4), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(dp, hp.data(), 8 * (n + 4), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(dk, hk.data(), 8 * (n + 4), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(db, hb.data(), 8 * (n + 3), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipMemcpyAsync(du, hu.data(), 8 * (n + 3), hipMemcpyHostToDevice, c->stream));
  Solve1DArgs a;
  memset(&a, 0, sizeof(a));
  a.c = dc;
  a.dxp = dp;
  a.dxb = db;
  a.u = du;
  a.k = dk;
  a.flux = df;
  a.dt = dt;
  a.bvalue[0] = bvalue[0];
  a.bvalue[1] = bvalue[1];
  a.bcode[0] = bcode[0];
  a.bcode[1] = bcode[1];
  a.n = n;
  a.limiter = limiter;
  hipError_t e = solve_1d(a, c->stream);
  if (e != hipSuccess) {
    dfree(d);
    return fail(c, FALL3D_ADS_EHIP, "solve_1d: %s", hipGetErrorString(e));
  }
  double fl[2];
  HIPCHK(c, hipMemcpyAsync(hc.data(), dc, 8 * (n + 4), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(fl, df, 16, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  dfree(d);
  for (int i = 0; i < n + 4; ++i) {
    if (c->kind == 8)
      ((double *)cl)[i] = hc[i];
    else
      ((float *)cl)[i] = (float)hc[i];
  }
  flux[0] = fl[0];
  flux[1] = fl[1];
  return FALL3D_ADS_OK;
}

// ---- resident group ---------------------------------------------------------------

int fall3d_ads_set_meteo(fall3d_ads_ctx *c, const void *u1, const void *u2, const void *v1, const void *v2,
                         const void *w1, const void *w2, const void *k1, const void *k2, const void *k3,
                         const void *vs) {
  if (!c) return FALL3D_ADS_EINVAL;
  if (!u1 || !v1 || !w1 || !k1 || !k2 || !k3 || !vs) return fail(c, FALL3D_ADS_EINVAL, "null meteo array");
  int rc;
  if ((rc = up_u(c, u1, c->u1, c->u1d))) return rc;
  if ((rc = up_u(c, u2 ? u2 : u1, c->u2, c->u2d))) return rc;
  if ((rc = up_v(c, v1, c->v1, c->v1d))) return rc;
  if ((rc = up_v(c, v2 ? v2 : v1, c->v2, c->v2d))) return rc;
  if ((rc = up_w(c, w1, c->w1, c->w1d))) return rc;
  if ((rc = up_w(c, w2 ? w2 : w1, c->w2, c->w2d))) return rc;
  if ((rc = up_k1(c, k1))) return rc;
  if ((rc = up_k2(c, k2))) return rc;
  if ((rc = up_k3(c, k3))) return rc;
  if ((rc = up_vs(c, vs))) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->meteo_ok = true;
  return FALL3D_ADS_OK;
}

int fall3d_ads_set_source_metric(fall3d_ads_ctx *c, const void *s, const void *Hm1_p, const void *Hm2_p,
                                 const void *Hm3_p) {
  if (!c) return FALL3D_ADS_EINVAL;
  int rc;
  if ((rc = need_grid(c))) return rc;
  dfree(c->src_idx);
  dfree(c->src_rate);
  dfree(c->src_meas);
  dfree(c->src_part);
  dfree(c->src_s);
  dfree(c->src_vol);
  c->src_idx = nullptr;
  c->src_rate = c->src_meas = c->src_part = c->src_s = c->src_vol = nullptr;
  c->part[0] = nullptr;
  c->nsrc = 0;
  c->version++;
  if (!s) return FALL3D_ADS_OK;
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  // map factors Hm1_p(1:ny), Hm2_p(-1:ny+2), Hm3_p(1:nx,1:ny); NULL = 1
  auto hm1 = [&](int j) { return Hm1_p ? host_val(Hm1_p, c->kind, (size_t)(j - 1)) : 1.0; };
  auto hm2 = [&](int j) { return Hm2_p ? host_val(Hm2_p, c->kind, (size_t)(j + 1)) : 1.0; };
  auto hm3 = [&](int i, int j) {
    return Hm3_p ? host_val(Hm3_p, c->kind, (size_t)(i - 1) + (size_t)nx * (size_t)(j - 1)) : 1.0;
  };
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i)
      if (!(hm1(j) * hm2(j) * hm3(i, j) > 0.0))
        return fail(c, FALL3D_ADS_EINVAL, "non-positive map factor product at cell (%d,%d)", i, j);
  std::vector<size_t> idx;  // element offsets in the state (c64 or chi/clo)
  std::vector<double> rate, meas, sval, volv;
  const int px = is_fp64(c) ? nx + 4 : state_pitch(nx);
  // Sources/mod_F3D.F90:1456-1473: c += dt*s/vol, vol = dX1_p dX2_p dX3_p Hm1 Hm2 Hm3
  for (int b = 0; b < c->nb; ++b)
    for (int k = 1; k <= nz; ++k)
      for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) {
          const size_t si = (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * ((size_t)(k - 1) + (size_t)nz * b));
          const double v = host_val(s, c->kind, si);
          if (v == 0.0) continue;
          const double vol = c->hdxp64[0][i + 1] * c->hdxp64[1][j + 1] * c->hdxp64[2][k + 1] *
                             (hm1(j) * hm2(j) * hm3(i, j));
          idx.push_back((size_t)(i + 1) + (size_t)px * ((size_t)(j + 1) + (size_t)(ny + 4) * ((size_t)(k + 1) + (size_t)(nz + 4) * b)));
          rate.push_back(v / vol);
          sval.push_back(v);
          volv.push_back(vol);
          meas.push_back(c->hdxb64[0][i] * c->hdxb64[1][j] * c->hdxb64[2][k]);
        }
  c->nsrc = (int)idx.size();
  if (c->nsrc == 0) return FALL3D_ADS_OK;
  HIPCHK(c, dalloc(&c->src_idx, idx.size()));
  HIPCHK(c, dalloc(&c->src_rate, idx.size()));
  HIPCHK(c, dalloc(&c->src_meas, idx.size()));
  const size_t nblk = (idx.size() + 255) / 256;
  HIPCHK(c, dalloc(&c->src_part, nblk * 4));
  c->part[0] = c->src_part;
  HIPCHK(c, hipMemcpy(c->src_idx, idx.data(), sizeof(size_t) * idx.size(), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->src_rate, rate.data(), sizeof(double) * idx.size(), hipMemcpyHostToDevice));
  HIPCHK(c, hipMemcpy(c->src_meas, meas.data(), sizeof(double) * idx.size(), hipMemcpyHostToDevice));
  if (is_fp64(c)) {
    // the FP64 state forms dt*s/vol as the reference does (not dt*(s/vol))
    HIPCHK(c, dalloc(&c->src_s, idx.size()));
    HIPCHK(c, dalloc(&c->src_vol, idx.size()));
    HIPCHK(c, hipMemcpy(c->src_s, sval.data(), sizeof(double) * idx.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->src_vol, volv.data(), sizeof(double) * idx.size(), hipMemcpyHostToDevice));
  }
  return FALL3D_ADS_OK;
}

int fall3d_ads_set_source(fall3d_ads_ctx *c, const void *s) {
  return fall3d_ads_set_source_metric(c, s, nullptr, nullptr, nullptr);
}

int fall3d_ads_set_state(fall3d_ads_ctx *c, const void *cst, const void *acum) {
  if (!c) return FALL3D_ADS_EINVAL;
  int rc;
  if (cst) {
    if ((rc = up_state(c, cst))) return rc;
  } else if (is_fp64(c)) {  // zero field, no host transfer
    HIPCHK(c, hipMemsetAsync(c->c64, 0, c->ncell_c * sizeof(double), c->stream));
  } else {
    HIPCHK(c, hipMemsetAsync(c->chi, 0, c->ncell_s * sizeof(float), c->stream));
    if (c->comp) HIPCHK(c, hipMemsetAsync(c->clo, 0, c->ncell_s * sizeof(float), c->stream));
  }
  if (acum) {
    if ((rc = up_acum(c, acum))) return rc;
  } else {
    HIPCHK(c, hipMemsetAsync(c->acum, 0, c->nacum * sizeof(double), c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FALL3D_ADS_OK;
}

int fall3d_ads_get_state(fall3d_ads_ctx *c, void *cst, void *acum) {
  if (!c) return FALL3D_ADS_EINVAL;
  int rc;
  if (cst && (rc = down_state(c, cst))) return rc;
  if (acum && (rc = down_acum(c, acum))) return rc;
  return FALL3D_ADS_OK;
}

int fall3d_ads_get_time_step(fall3d_ads_ctx *c, double stime, int criterion, double cfl, double *dt) {
  if (!c || !dt) return FALL3D_ADS_EINVAL;
  int rc;
  if ((rc = need_grid(c))) return rc;
  if (!c->meteo_ok) return fail(c, FALL3D_ADS_ESTATE, "fall3d_ads_set_meteo has not been called");
  if (criterion != 1 && criterion != 2) return fail(c, FALL3D_ADS_EINVAL, "invalid CFL criterion %d", criterion);
  CflArgs a;
  memset(&a, 0, sizeof(a));
  a.nx = c->nx;
  a.ny = c->ny;
  a.nz = c->nz;
  a.nb = c->nb;
  a.criterion = criterion;
  a.stime = stime;
  a.cfl = cfl;
  if (is_fp64(c)) {  // the FP64 state's own fields: the reference's dt in FP64
    a.fp64 = 1;
    a.u1 = c->u1d; a.u2 = c->u2d; a.v1 = c->v1d; a.v2 = c->v2d; a.w1 = c->w1d; a.w2 = c->w2d;
    a.k1 = c->k1d; a.k2 = c->k2d; a.k3 = c->k3d; a.vs = c->vsd;
  } else {
    a.u1 = c->u1; a.u2 = c->u2; a.v1 = c->v1; a.v2 = c->v2; a.w1 = c->w1; a.w2 = c->w2;
    a.k1 = c->k1; a.k2 = c->k2; a.k3 = c->k3; a.vs = c->vs;
  }
  a.dx1p = c->dxp64[0];
  a.dx2p = c->dxp64[1];
  a.dx3p = c->dxp64[2];
  HIPCHK(c, cfl_time_step(a, c->red_part, c->red_cap - 1, c->red_part + (c->red_cap - 1), c->stream));
  HIPCHK(c, hipMemcpyAsync(dt, c->red_part + (c->red_cap - 1), sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FALL3D_ADS_OK;
}

int fall3d_ads_step(fall3d_ads_ctx *c, double dt, double stime, int iiter) {
  if (!c) return FALL3D_ADS_EINVAL;
  return step_impl(c, dt, stime, iiter, nullptr);
}

int fall3d_ads_run(fall3d_ads_ctx *c, double dt, double stime, int iiter0, int nsteps) {
  if (!c) return FALL3D_ADS_EINVAL;
  int rc;
  // Launch-bound inner loop: replay a captured (even, odd) step pair as one
  // hipGraph. Per-kernel timing needs events between kernels, so timed steps
  // launch directly: all of them when every step is timed (no graph), one in
  // prof_every otherwise (the pairs in between are replays).
  const int every = c->prof_every;
  const bool sampling = every > 1;
  const bool graph = c->use_graph && every != 1 && nsteps >= 4;
  if (graph) {
    // the dt/dxb tables are refreshed outside the graph (they may be stale
    // after fall3d_ads_set_grid even when the graph itself is reused)
    if ((rc = ensure_dt(c, dt))) return rc;
    if (!c->gexec || c->g_dt != dt || c->g_stime != stime || c->g_version != c->version) {
      if (c->gexec) {
        (void)hipGraphExecDestroy(c->gexec);
        c->gexec = nullptr;
      }
      const bool prof = c->prof;
      c->prof = false;  // no events inside the graph
      HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int r0 = fall3d_ads_step(c, dt, stime, 0);
      int r1 = r0 ? r0 : fall3d_ads_step(c, dt, stime, 1);
      hipGraph_t gr = nullptr;
      hipError_t e = hipStreamEndCapture(c->stream, &gr);
      c->prof = prof;
      if (r1) {
        if (gr) (void)hipGraphDestroy(gr);
        return r1;
      }
      if (e != hipSuccess) return fail(c, FALL3D_ADS_EHIP, "step capture: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&c->gexec, gr, nullptr, nullptr, 0);
      (void)hipGraphDestroy(gr);
      if (e != hipSuccess) {
        c->gexec = nullptr;
        return fail(c, FALL3D_ADS_EHIP, "graph instantiate: %s", hipGetErrorString(e));
      }
      c->g_dt = dt;
      c->g_stime = stime;
      c->g_version = c->version;
    }
  }
  int it = iiter0, left = nsteps;
  while (left > 0) {
    const int k = it - iiter0;  // timed steps: k = 0, every, 2 every, ...
    const bool t0 = sampling && k % every == 0, t1 = sampling && (k + 1) % every == 0;
    if (graph && left >= 2 && it % 2 == 0 && !t0 && !t1) {
      HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
      it += 2;
      left -= 2;
      continue;
    }
    if (sampling) c->prof = t0;
    rc = fall3d_ads_step(c, dt, stime, it);
    if (sampling) c->prof = true;
    if (rc) return rc;
    ++it;
    --left;
  }
  return FALL3D_ADS_OK;
}

int fall3d_ads_step_noalloc(fall3d_ads_ctx *c, double dt, double stime, int iiter) {
  if (!c) return FALL3D_ADS_EINVAL;
  if (c->learned)
    return fail(c, FALL3D_ADS_EINVAL, "the allocation-free step runs plain KT only (learned correction is set)");
  if (is_fp64(c)) return fail(c, FALL3D_ADS_EINVAL, "the allocation-free step runs the FP32 states only");
  if (c->prof || c->prof_every)
    return fail(c, FALL3D_ADS_EINVAL, "the allocation-free step does not record profiling events");
  double *scr = nullptr;
  int rc;
  if ((rc = noalloc_scratch(c, &scr))) return rc;
  // partial rows (4 doubles) of the source and of each sweep's workgroups
  size_t rows[4];
  rows[0] = (size_t)(c->nsrc + 255) / 256;
  for (int d = 0; d < 3; ++d) rows[d + 1] = (d == 0) ? (size_t)c->tl[0].gx * c->tl[0].groups : (size_t)c->tl[d].gx;
  // the sweeps' rows twice: their partial sums and their rounding residuals
  const size_t need = 4 * (rows[0] + 2 * (rows[1] + rows[2] + rows[3]));
  if (need > kNoAllocScratchDoubles)
    return fail(c, FALL3D_ADS_EINVAL,
                "the step needs %zu doubles of partial sums; the allocation-free entry point holds %zu "
                "(use fall3d_ads_step, or fewer workgroups: more bins per workgroup)",
                need, kNoAllocScratchDoubles);
  double *part[7];
  size_t off = 0;
  for (int q = 0; q < 7; ++q) {
    part[q] = scr + off;
    off += 4 * rows[q < 4 ? q : q - 3];
  }
  return step_impl(c, dt, stime, iiter, part);
}

int fall3d_ads_copy_probe(const void *src, void *dst, size_t n, void *stream) {
  if ((!src || !dst) && n) return FALL3D_ADS_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (copy_probe((const float *)src, (float *)dst, n, st) != hipSuccess) return FALL3D_ADS_EHIP;
  return hipStreamSynchronize(st) == hipSuccess ? FALL3D_ADS_OK : FALL3D_ADS_EHIP;
}

int fall3d_ads_copy_probe64(const void *src, void *dst, size_t n, void *stream) {
  if ((!src || !dst) && n) return FALL3D_ADS_EINVAL;
  hipStream_t st = (hipStream_t)stream;
  if (copy_probe64((const double *)src, (double *)dst, n, st) != hipSuccess) return FALL3D_ADS_EHIP;
  return hipStreamSynchronize(st) == hipSuccess ? FALL3D_ADS_OK : FALL3D_ADS_EHIP;
}

int fall3d_ads_set_boundary(fall3d_ads_ctx *c, int mode) {
  if (!c) return FALL3D_ADS_EINVAL;
  if (mode < FALL3D_ADS_BC_FREEFLOW || mode > FALL3D_ADS_BC_PERIODIC)
    return fail(c, FALL3D_ADS_EINVAL, "invalid boundary condition %d", mode);
  c->bc = mode;
  c->version++;  // captured steps hold the old rule in their arguments
  return FALL3D_ADS_OK;
}

int fall3d_ads_set_graph(fall3d_ads_ctx *c, int on) {
  if (!c) return FALL3D_ADS_EINVAL;
  c->use_graph = on != 0;
  return FALL3D_ADS_OK;
}

int fall3d_ads_get_fluxes(fall3d_ads_ctx *c, double flux[6]) {
  if (!c || !flux) return FALL3D_ADS_EINVAL;
  return get_acc(c, ACC_FLUX, flux, 6);
}

int fall3d_ads_set_fluxes(fall3d_ads_ctx *c, const double flux[6]) {
  if (!c || !flux) return FALL3D_ADS_EINVAL;
  int rc = set_acc(c, ACC_FLUX, flux, 6);
  if (rc) return rc;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FALL3D_ADS_OK;
}

int fall3d_ads_get_budget(fall3d_ads_ctx *c, double out[FALL3D_ADS_BUDGET_LEN]) {
  if (!c || !out) return FALL3D_ADS_EINVAL;
  int rc;
  if ((rc = need_grid(c))) return rc;
  double *o = c->red_part + (c->red_cap - 2);
  // mass with mass-point spacings and with cell sizes, in one pass
  const MassSpacings sp = {c->dxp64[0], c->dxp64[1], c->dxp64[2], c->dxb64[0], c->dxb64[1], c->dxb64[2]};
  HIPCHK(c, mass(c->chi, c->comp ? c->clo : nullptr, c->c64, c->nx, c->ny, c->nz, c->nb, sp, c->red_part,
                 c->red_cap - 2, o, c->stream));
  double m2[2] = {0.0, 0.0};
  HIPCHK(c, hipMemcpyAsync(m2, o, sizeof(m2), hipMemcpyDeviceToHost, c->stream));
  double acc[ACC_LEN];
  HIPCHK(c, hipMemcpyAsync(acc, c->acc, sizeof(acc), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  out[FALL3D_ADS_BUDGET_MASS_VOLUME] = m2[0];
  out[FALL3D_ADS_BUDGET_MASS_CELLS] = m2[1];
  out[FALL3D_ADS_BUDGET_GROUND] = fabs(acc[ACC_FLUX + 4]);
  out[FALL3D_ADS_BUDGET_LATERAL] =
      fabs(acc[ACC_FLUX + 0]) + fabs(acc[ACC_FLUX + 1]) + fabs(acc[ACC_FLUX + 2]) + fabs(acc[ACC_FLUX + 3]) +
      fabs(acc[ACC_FLUX + 5]);
  out[FALL3D_ADS_BUDGET_INJECTED] = acc[ACC_INJECTED];
  for (int s = 0; s < 6; ++s) out[FALL3D_ADS_BUDGET_LEDGER_W + s] = acc[ACC_LEDGER + s];
  out[FALL3D_ADS_BUDGET_RESIDUAL] = acc[ACC_RESIDUAL + 0] + acc[ACC_RESIDUAL + 1] + acc[ACC_RESIDUAL + 2];
  return FALL3D_ADS_OK;
}

// ---- restart snapshot ------------------------------------------------------
// Layout (little-endian): 64-byte header, then the accumulators (ACC_LEN
// doubles), acum (FP64), the FP32 state words and, for the compensated
// state, its low words -- everything the device holds that a resumed run
// reads, so that the resumed steps are bitwise those of an unbroken run.
namespace {
struct SnapHeader {
  char magic[8];       // "F3DADSR1"
  uint32_t version;    // 1
  int32_t nx, ny, nz, nb, comp, limiter, tm;
  double time;         // simulated time of the snapshot (caller's value)
  int64_t iiter;       // next iteration (caller's value)
  uint64_t acc_len;
};
static_assert(sizeof(SnapHeader) == 64, "snapshot header is 64 bytes");
constexpr char kSnapMagic[8] = {'F', '3', 'D', 'A', 'D', 'S', 'R', '1'};

// bytes of the state words: FP32 hi (+ lo), or FP64
size_t state_bytes(const fall3d_ads_ctx *c) {
  return is_fp64(c) ? sizeof(double) * c->ncell_c : sizeof(float) * c->ncell_s * (c->comp ? 2 : 1);
}

size_t snapshot_bytes(const fall3d_ads_ctx *c) {
  return sizeof(SnapHeader) + sizeof(double) * ((size_t)ACC_LEN + c->nacum) + state_bytes(c);
}
}  // namespace

int fall3d_ads_snapshot(fall3d_ads_ctx *c, void *buf, size_t *bytes, double time, long long iiter) {
  if (!c || !bytes) return FALL3D_ADS_EINVAL;
  const size_t need = snapshot_bytes(c);
  if (!buf) {
    *bytes = need;
    return FALL3D_ADS_OK;
  }
  if (*bytes < need) return fail(c, FALL3D_ADS_EINVAL, "snapshot buffer holds %zu bytes, %zu needed", *bytes, need);
  SnapHeader h;
  memset(&h, 0, sizeof(h));
  memcpy(h.magic, kSnapMagic, 8);
  h.version = 1;
  h.nx = c->nx;
  h.ny = c->ny;
  h.nz = c->nz;
  h.nb = c->nb;
  h.comp = c->comp;
  h.limiter = c->limiter;
  h.tm = c->tm;
  h.time = time;
  h.iiter = iiter;
  h.acc_len = ACC_LEN;
  char *p = (char *)buf;
  memcpy(p, &h, sizeof(h));
  p += sizeof(h);
  HIPCHK(c, hipMemcpyAsync(p, c->acc, sizeof(double) * ACC_LEN, hipMemcpyDeviceToHost, c->stream));
  p += sizeof(double) * ACC_LEN;
  HIPCHK(c, hipMemcpyAsync(p, c->acum, sizeof(double) * c->nacum, hipMemcpyDeviceToHost, c->stream));
  p += sizeof(double) * c->nacum;
  if (is_fp64(c)) {
    HIPCHK(c, hipMemcpyAsync(p, c->c64, state_bytes(c), hipMemcpyDeviceToHost, c->stream));
  } else {
    HIPCHK(c, hipMemcpyAsync(p, c->chi, sizeof(float) * c->ncell_s, hipMemcpyDeviceToHost, c->stream));
    p += sizeof(float) * c->ncell_s;
    if (c->comp) HIPCHK(c, hipMemcpyAsync(p, c->clo, sizeof(float) * c->ncell_s, hipMemcpyDeviceToHost, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *bytes = need;
  return FALL3D_ADS_OK;
}

int fall3d_ads_restore(fall3d_ads_ctx *c, const void *buf, size_t bytes, double *time, long long *iiter) {
  if (!c || !buf) return FALL3D_ADS_EINVAL;
  SnapHeader h;
  if (bytes < sizeof(h)) return fail(c, FALL3D_ADS_EINVAL, "snapshot of %zu bytes has no header", bytes);
  memcpy(&h, buf, sizeof(h));
  if (memcmp(h.magic, kSnapMagic, 8) != 0 || h.version != 1)
    return fail(c, FALL3D_ADS_EINVAL, "not an ADS snapshot (magic or version)");
  if (h.nx != c->nx || h.ny != c->ny || h.nz != c->nz || h.nb != c->nb)
    return fail(c, FALL3D_ADS_EINVAL, "snapshot of a %d x %d x %d x %d grid, context is %d x %d x %d x %d", h.nx, h.ny,
                h.nz, h.nb, c->nx, c->ny, c->nz, c->nb);
  if (h.comp != c->comp) return fail(c, FALL3D_ADS_EINVAL, "snapshot state mode %d, context %d", h.comp, c->comp);
  if (h.acc_len != (uint64_t)ACC_LEN) return fail(c, FALL3D_ADS_EINVAL, "snapshot accumulator layout differs");
  const size_t need = snapshot_bytes(c);
  if (bytes < need) return fail(c, FALL3D_ADS_EINVAL, "snapshot of %zu bytes, %zu expected", bytes, need);
  const char *p = (const char *)buf + sizeof(h);
  HIPCHK(c, hipMemcpyAsync(c->acc, p, sizeof(double) * ACC_LEN, hipMemcpyHostToDevice, c->stream));
  p += sizeof(double) * ACC_LEN;
  HIPCHK(c, hipMemcpyAsync(c->acum, p, sizeof(double) * c->nacum, hipMemcpyHostToDevice, c->stream));
  p += sizeof(double) * c->nacum;
  if (is_fp64(c)) {
    HIPCHK(c, hipMemcpyAsync(c->c64, p, state_bytes(c), hipMemcpyHostToDevice, c->stream));
  } else {
    HIPCHK(c, hipMemcpyAsync(c->chi, p, sizeof(float) * c->ncell_s, hipMemcpyHostToDevice, c->stream));
    p += sizeof(float) * c->ncell_s;
    if (c->comp) HIPCHK(c, hipMemcpyAsync(c->clo, p, sizeof(float) * c->ncell_s, hipMemcpyHostToDevice, c->stream));
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (time) *time = h.time;
  if (iiter) *iiter = h.iiter;
  return FALL3D_ADS_OK;
}

int fall3d_ads_check_state(fall3d_ads_ctx *c, long long counts[2], double *min_value) {
  if (!c || !counts) return FALL3D_ADS_EINVAL;
  double *o = c->red_part + (c->red_cap - 3);
  HIPCHK(c, health(c->chi, c->comp ? c->clo : nullptr, c->c64, c->nx, c->ny, c->nz, c->nb, c->red_part,
                   c->red_cap - 3, o, c->stream));
  double h[3];
  HIPCHK(c, hipMemcpyAsync(h, o, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  counts[0] = (long long)h[0];
  counts[1] = (long long)h[1];
  if (min_value) *min_value = h[2];
  return FALL3D_ADS_OK;
}

int fall3d_ads_set_stream(fall3d_ads_ctx *c, void *stream) {
  if (!c) return FALL3D_ADS_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (stream) {
    if (c->own_stream) (void)hipStreamDestroy(c->stream);
    c->stream = (hipStream_t)stream;
    c->own_stream = false;
  } else if (!c->own_stream) {
    HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = true;
  }
  return FALL3D_ADS_OK;
}

int fall3d_ads_synchronize(fall3d_ads_ctx *c) {
  if (!c) return FALL3D_ADS_EINVAL;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return FALL3D_ADS_OK;
}

// Autotuner of the tile shapes: every candidate of each sweep is timed
// (reps launches between HIP events on the context stream, after one
// warm-up launch) on the context's own fields; the fastest is kept. The
// sweeps change the state, so state, deposit and accumulators are saved
// before and restored after.
int fall3d_ads_autotune(fall3d_ads_ctx *c, double dt, double stime, int reps, int tiling[3], double best_ms[3]) {
  if (!c) return FALL3D_ADS_EINVAL;
  int rc;
  if ((rc = need_grid(c))) return rc;
  if (!c->meteo_ok) return fail(c, FALL3D_ADS_ESTATE, "fall3d_ads_set_meteo has not been called");
  if (is_fp64(c)) return fail(c, FALL3D_ADS_EINVAL, "the FP64 state has no tile shapes to tune");
  if (reps < 1) reps = 3;
  if ((rc = ensure_dt(c, dt))) return rc;
  // save
  void *src[4] = {c->chi, c->comp ? c->clo : nullptr, c->acum, c->acc};
  const size_t bytes[4] = {c->ncell_s * sizeof(float), c->ncell_s * sizeof(float), c->nacum * sizeof(double),
                           (size_t)ACC_LEN * sizeof(double)};
  void *save[4] = {nullptr, nullptr, nullptr, nullptr};
  for (int i = 0; i < 4; ++i) {
    if (!src[i]) continue;
    HIPCHK(c, hipMalloc(&save[i], bytes[i]));
    HIPCHK(c, hipMemcpyAsync(save[i], src[i], bytes[i], hipMemcpyDeviceToDevice, c->stream));
  }
  hipEvent_t ea = nullptr, eb = nullptr;
  HIPCHK(c, hipEventCreate(&ea));
  HIPCHK(c, hipEventCreate(&eb));
  const bool prof = c->prof;
  c->prof = false;
  const int oc[3] = {c->user_cpt[0], c->user_cpt[1], c->user_cpt[2]};
  const int ob[3] = {c->user_bpg[0], c->user_bpg[1], c->user_bpg[2]};
  double best[3] = {1e30, 1e30, 1e30};
  int pick[3] = {0, 0, 0};
  for (int d = 0; d < 3 && !rc; ++d) {
    // x: cells per lane; y/z: bins per group g and bins per thread b as g + 16 (b - 1)
    std::vector<int> cand;
    if (d == 0) {
      for (int i = 0; i < kSweepCptCount; ++i)
        if (c->nx / kSweepCpt[i] + 1 <= 256) cand.push_back(kSweepCpt[i]);  // up to 4 waves per line
    } else {
      for (int b = 1; b <= (c->comp ? 1 : 2); ++b)
        for (int g = b; g <= std::min(c->nb, 8); ++g) {
          const int groups = (c->nb + g - 1) / g;
          if ((c->nb + groups - 1) / groups == g) cand.push_back(g + 16 * (b - 1));
        }
    }
    for (int v : cand) {
      if (d == 0) {
        c->user_cpt[0] = v;
      } else {
        c->user_bpg[d] = v & 15;
        c->user_cpt[d] = (v >> 4) + 1;
      }
      if ((rc = refresh_tiling(c))) break;
      const float *v1 = d == 0 ? c->u1 : (d == 1 ? c->v1 : c->w1);
      const float *v2 = d == 0 ? c->u2 : (d == 1 ? c->v2 : c->w2);
      const float *kd = d == 0 ? c->k1 : (d == 1 ? c->k2 : c->k3);
      // a candidate costs its sweep plus the reduction of its partials
      auto sweep_and_reduce = [&]() -> int {
        SweepOut so;
        int r = run_sweep(c, d, dt, stime, v1, v2, kd, 1, &so);
        if (r) return r;
        ReduceJob j;
        memset(&j, 0, sizeof(j));
        add_sweep_jobs(j, c, d, so.nblocks);
        return run_finalize(c, j);
      };
      if ((rc = sweep_and_reduce())) break;
      (void)hipEventRecord(ea, c->stream);
      for (int r = 0; r < reps && !rc; ++r) rc = sweep_and_reduce();
      (void)hipEventRecord(eb, c->stream);
      if (rc) break;
      (void)hipEventSynchronize(eb);
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, ea, eb);
      if (ms / reps < best[d]) {
        best[d] = ms / reps;
        pick[d] = v;
      }
    }
    if (d == 0) {
      c->user_cpt[0] = pick[0] ? pick[0] : oc[0];
    } else {
      c->user_bpg[d] = pick[d] ? (pick[d] & 15) : ob[d];
      c->user_cpt[d] = pick[d] ? (pick[d] >> 4) + 1 : oc[d];
    }
  }
  c->prof = prof;
  (void)hipEventDestroy(ea);
  (void)hipEventDestroy(eb);
  // restore
  for (int i = 0; i < 4; ++i) {
    if (!save[i]) continue;
    if (!rc) {
      const hipError_t e = hipMemcpyAsync(src[i], save[i], bytes[i], hipMemcpyDeviceToDevice, c->stream);
      if (e != hipSuccess) rc = fail(c, FALL3D_ADS_EHIP, "autotune restore: %s", hipGetErrorString(e));
    }
  }
  (void)hipStreamSynchronize(c->stream);
  for (int i = 0; i < 4; ++i)
    if (save[i]) (void)hipFree(save[i]);
  if (rc) {
    for (int d = 0; d < 3; ++d) {
      c->user_cpt[d] = oc[d];
      c->user_bpg[d] = ob[d];
    }
    (void)refresh_tiling(c);
    return rc;
  }
  if ((rc = refresh_tiling(c))) return rc;
  if (tiling) {
    tiling[0] = c->tl[0].cpt;
    tiling[1] = c->tl[1].bpg;
    tiling[2] = c->tl[2].bpg;
  }
  if (best_ms)
    for (int d = 0; d < 3; ++d) best_ms[d] = best[d];
  return FALL3D_ADS_OK;
}

int fall3d_ads_get_tiling(fall3d_ads_ctx *c, int dir, int *cells_per_thread, int *bins_per_group) {
  if (!c || dir < 0 || dir > 2) return FALL3D_ADS_EINVAL;
  if (cells_per_thread) *cells_per_thread = c->tl[dir].cpt;
  if (bins_per_group) *bins_per_group = c->tl[dir].bpg;
  return FALL3D_ADS_OK;
}

int fall3d_ads_set_tiling(fall3d_ads_ctx *c, int dir, int cells_per_thread, int columns, int bins_per_group) {
  if (!c || dir < 0 || dir > 2) return FALL3D_ADS_EINVAL;
  if (is_fp64(c)) return fail(c, FALL3D_ADS_EINVAL, "the FP64 state has no tile shapes to set");
  if (dir == 0 && cells_per_thread && !sweep_cpt_supported(cells_per_thread))
    return fail(c, FALL3D_ADS_EINVAL, "unsupported cells per thread %d", cells_per_thread);
  // y/z: bins per thread, 1 or 2 (2 is not built for the compensated state)
  if (dir > 0 && (cells_per_thread < 0 || cells_per_thread > 2 || (cells_per_thread == 2 && c->comp)))
    return fail(c, FALL3D_ADS_EINVAL, "unsupported bins per thread %d", cells_per_thread);
  if (columns && (columns < 1 || columns > 64)) return fail(c, FALL3D_ADS_EINVAL, "columns %d out of range", columns);
  const int oc = c->user_cpt[dir], ot = c->user_tx[dir], ob = c->user_bpg[dir];
  c->user_cpt[dir] = cells_per_thread;
  c->user_tx[dir] = columns;
  c->user_bpg[dir] = bins_per_group;
  int rc = refresh_tiling(c);
  if (rc) {
    c->user_cpt[dir] = oc;
    c->user_tx[dir] = ot;
    c->user_bpg[dir] = ob;
    (void)refresh_tiling(c);
  }
  return rc;
}

int fall3d_ads_set_profiling(fall3d_ads_ctx *c, int on) {
  if (!c) return FALL3D_ADS_EINVAL;
  drain_events(c);
  if (on < 0) return fail(c, FALL3D_ADS_EINVAL, "profiling interval must be >= 0");
  c->prof = on != 0;
  c->prof_every = on;
  for (int k = 0; k < KID_COUNT; ++k) {
    c->kt_ms[k] = 0.0;
    c->kt_n[k] = 0;
  }
  return FALL3D_ADS_OK;
}

int fall3d_ads_get_kernel_time(fall3d_ads_ctx *c, int kernel, double *total_ms, long long *launches) {
  if (!c || kernel < 0 || kernel >= KID_COUNT || !total_ms || !launches) return FALL3D_ADS_EINVAL;
  drain_events(c);
  *total_ms = c->kt_ms[kernel];
  *launches = c->kt_n[kernel];
  return FALL3D_ADS_OK;
}

int fall3d_ads_sweep_bytes(fall3d_ads_ctx *c, int dir, double *bytes) {
  if (!c || !bytes || dir < 0 || dir > 2) return FALL3D_ADS_EINVAL;
  const double cells = (double)c->nx * c->ny * c->nz;
  const double st = c->comp ? 16.0 : 8.0;      // state read + write per cell and bin (FP32 pair or FP64)
  const double fw = is_fp64(c) ? 8.0 : 4.0;     // bytes of a field value
  // coefficient fields read once per sweep: two wind snapshots + diffusion
  double b = cells * (c->nb * st + 3.0 * fw);
  if (dir == 2) b += cells * c->nb * fw;  // settling velocity per bin
  *bytes = b;
  return FALL3D_ADS_OK;
}

}  // extern "C"
