// ads_ensemble.hip -- ensemble members of different sizes advanced together:
// one launch per sweep (and one for the sources, one for the reductions)
// covers every member, and a whole run replays from one captured hipGraph.
//
// FALL3D runs ensemble members as independent copies of task_Fall3d, one
// per communicator (Sources/mod_Ensemble.F90, Sources/mod_Parallel.F90:
// 188-196); each member calls the same ADS_solve_along_* on its own grid.
// Here each member keeps its own context (state, meteo, grid, sources,
// accumulators, time step), created and filled through the single-domain
// entry points; the ensemble only adds the batched launches. A member's
// result is bit-identical to running its context alone with the same
// tiling, because the batched kernels run the same kernel bodies on the same
// per-member arguments.
//
// The autotuner times the candidate tile shapes of each sweep on the actual
// members and keeps the fastest (the state is restored afterwards).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "ads_context.h"

using namespace f3d;
using namespace f3d::host;

struct fall3d_ads_ensemble {
  int nmem = 0, nb = 0, limiter = 1, tm = 2, kind = 8, comp = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err = "no error";
  std::vector<fall3d_ads_ctx *> mem;

  // common tile shape: cells per lane of the x sweep, bins per group of y/z
  int x_cpt = 0, bpg[3] = {0, 0, 0};
  int bpt[3] = {0, 0, 0};  // y/z bins per thread (0: the members' default)

  // per axis: member arguments and workgroup prefix of the batched sweep
  SweepArgs *dargs[3] = {nullptr, nullptr, nullptr};
  int *dbstart[3] = {nullptr, nullptr, nullptr};
  int nblocks[3] = {0, 0, 0};
  int block[3] = {0, 0, 0};
  // sources
  SourceMember *dsrc = nullptr;
  int *dsrc_bstart = nullptr;
  int src_blocks = 0;
  // reductions
  BatchJob *djobs = nullptr;
  int njobs = 0;
  // device tables are valid for these time steps and this version
  std::vector<double> tab_dt, tab_stime;
  std::vector<int> tab_version;
  bool tab_ok = false;

  // graph of a whole run
  hipGraphExec_t gexec = nullptr;
  int g_steps = 0, g_iiter_parity = -1;
  std::vector<double> g_dt, g_stime;
  std::vector<int> g_version;
  bool use_graph = true;
};

namespace {

int efail(fall3d_ads_ensemble *e, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (e) e->err = buf;
  return code;
}

#define EHIPCHK(ens, expr)                                                                           \
  do {                                                                                               \
    hipError_t e_ = (expr);                                                                          \
    if (e_ != hipSuccess) return efail((ens), FALL3D_ADS_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// member m failed: copy its message
int mfail(fall3d_ads_ensemble *e, int m, int rc) {
  return efail(e, rc, "member %d: %s", m, fall3d_ads_last_error(e->mem[m]));
}

void drop_graph(fall3d_ads_ensemble *e) {
  if (e->gexec) (void)hipGraphExecDestroy(e->gexec);
  e->gexec = nullptr;
  e->g_steps = 0;
}

// Apply the common tile shape to every member (their own launches then use
// it too, so a member run alone matches its batched result bit for bit).
int apply_tiling(fall3d_ads_ensemble *e) {
  for (int m = 0; m < e->nmem; ++m) {
    fall3d_ads_ctx *c = e->mem[m];
    c->user_cpt[0] = e->x_cpt;
    c->user_bpg[0] = c->nb;  // one bin group: the batch already has enough workgroups
    c->user_bpg[1] = e->bpg[1];
    c->user_bpg[2] = e->bpg[2];
    c->user_cpt[1] = e->bpt[1];
    c->user_cpt[2] = e->bpt[2];
    const int rc = refresh_tiling(c);
    if (rc) return mfail(e, m, rc);
  }
  e->tab_ok = false;
  drop_graph(e);
  return FALL3D_ADS_OK;
}

int members_ready(fall3d_ads_ensemble *e) {
  for (int m = 0; m < e->nmem; ++m) {
    fall3d_ads_ctx *c = e->mem[m];
    if (!c->grid_ok) return efail(e, FALL3D_ADS_ESTATE, "member %d: fall3d_ads_set_grid has not been called", m);
    if (!c->meteo_ok) return efail(e, FALL3D_ADS_ESTATE, "member %d: fall3d_ads_set_meteo has not been called", m);
  }
  return FALL3D_ADS_OK;
}

bool tables_current(const fall3d_ads_ensemble *e, const double *dt, const double *stime) {
  if (!e->tab_ok) return false;
  for (int m = 0; m < e->nmem; ++m) {
    if (e->tab_dt[m] != dt[m] || e->tab_stime[m] != stime[m] || e->tab_version[m] != e->mem[m]->version)
      return false;
  }
  return true;
}

// Build (on the host) and upload the member tables of the batched launches
// for these time steps. Stream-ordered uploads; not done under capture.
int build_tables(fall3d_ads_ensemble *e, const double *dt, const double *stime) {
  if (tables_current(e, dt, stime)) return FALL3D_ADS_OK;
  int rc;
  const int M = e->nmem;
  for (int d = 0; d < 3; ++d) {
    std::vector<SweepArgs> args(M);
    std::vector<int> bstart(M + 1, 0);
    for (int m = 0; m < M; ++m) {
      fall3d_ads_ctx *c = e->mem[m];
      if ((rc = ensure_dt(c, dt[m]))) return mfail(e, m, rc);
      make_sweep_args(c, d, dt[m], stime[m], 1, args[m]);
      const Tiling &t = c->tl[d];
      const int nbl = (d == 0) ? t.gx * t.groups : t.gx;
      bstart[m + 1] = bstart[m] + nbl;
      if (m == 0) e->block[d] = t.block;
      if (t.block != e->block[d]) return efail(e, FALL3D_ADS_EINVAL, "members disagree on the %c-sweep block", "xyz"[d]);
    }
    e->nblocks[d] = bstart[M];
    EHIPCHK(e, hipMemcpyAsync(e->dargs[d], args.data(), sizeof(SweepArgs) * M, hipMemcpyHostToDevice, e->stream));
    EHIPCHK(e, hipMemcpyAsync(e->dbstart[d], bstart.data(), sizeof(int) * (M + 1), hipMemcpyHostToDevice,
                              e->stream));
  }
  // sources
  {
    std::vector<SourceMember> tab(M);
    std::vector<int> bstart(M + 1, 0);
    for (int m = 0; m < M; ++m) {
      fall3d_ads_ctx *c = e->mem[m];
      SourceMember &s = tab[m];
      memset(&s, 0, sizeof(s));
      s.chi = c->chi;
      s.clo = c->comp ? c->clo : nullptr;
      s.idx = c->src_idx;
      s.rate_vol = c->src_rate;
      s.meas = c->src_meas;
      s.part = c->src_part;
      s.dt = dt[m];
      s.nent = c->nsrc;
      bstart[m + 1] = bstart[m] + (c->nsrc + 255) / 256;
    }
    e->src_blocks = bstart[M];
    EHIPCHK(e, hipMemcpyAsync(e->dsrc, tab.data(), sizeof(SourceMember) * M, hipMemcpyHostToDevice, e->stream));
    EHIPCHK(e, hipMemcpyAsync(e->dsrc_bstart, bstart.data(), sizeof(int) * (M + 1), hipMemcpyHostToDevice,
                              e->stream));
  }
  // reductions: injected mass, then 4 slots per sweep (reference flux
  // begin/end, ledger begin/end), per member
  {
    std::vector<BatchJob> jobs;
    for (int m = 0; m < M; ++m) {
      fall3d_ads_ctx *c = e->mem[m];
      if (c->nsrc > 0) {
        BatchJob jb;
        memset(&jb, 0, sizeof(jb));
        jb.part = c->src_part;
        jb.dst[0] = c->acc + ACC_INJECTED;
        jb.nblocks = (c->nsrc + 255) / 256;
        jb.nslot = 1;
        jobs.push_back(jb);
      }
      for (int d = 0; d < 3; ++d) {
        const Tiling &t = c->tl[d];
        BatchJob jb;
        jb.part = c->part[d + 1];
        jb.dst[0] = c->acc + ACC_FLUX + 2 * d;
        jb.dst[1] = c->acc + ACC_FLUX + 2 * d + 1;
        jb.dst[2] = c->acc + ACC_LEDGER + 2 * d;
        jb.dst[3] = c->acc + ACC_LEDGER + 2 * d + 1;
        jb.nblocks = (d == 0) ? t.gx * t.groups : t.gx;
        jb.nslot = 4;
        jobs.push_back(jb);
        BatchJob jr;
        memset(&jr, 0, sizeof(jr));
        jr.part = c->rpart[d + 1];
        jr.dst[0] = c->acc + ACC_RESIDUAL + d;
        jr.nblocks = jb.nblocks;
        jr.nslot = 1;
        jobs.push_back(jr);
      }
    }
    e->njobs = (int)jobs.size();
    if (e->djobs) (void)hipFree(e->djobs);
    e->djobs = nullptr;
    EHIPCHK(e, dalloc(&e->djobs, jobs.size()));
    EHIPCHK(e, hipMemcpyAsync(e->djobs, jobs.data(), sizeof(BatchJob) * jobs.size(), hipMemcpyHostToDevice,
                              e->stream));
  }
  // the host vectors above go out of scope: finish the copies first
  EHIPCHK(e, hipStreamSynchronize(e->stream));
  e->tab_dt.assign(dt, dt + M);
  e->tab_stime.assign(stime, stime + M);
  e->tab_version.resize(M);
  for (int m = 0; m < M; ++m) e->tab_version[m] = e->mem[m]->version;
  e->tab_ok = true;
  return FALL3D_ADS_OK;
}

int batch_sweep(fall3d_ads_ensemble *e, int d) {
  SweepBatch b;
  b.args = e->dargs[d];
  b.bstart = e->dbstart[d];
  b.nmem = e->nmem;
  const int cpt = (d == 0) ? e->x_cpt : e->mem[0]->tl[d].cpt;
  hipError_t r = launch_sweep_batch(d, cpt, e->limiter, e->comp != 0, e->tm == 1 ? 1 : 4, b, e->nblocks[d],
                                    dim3(e->block[d]), e->stream);
  if (r != hipSuccess) return efail(e, FALL3D_ADS_EHIP, "batched %c sweep: %s", "xyz"[d], hipGetErrorString(r));
  return FALL3D_ADS_OK;
}

// One F3D_time_step of every member (tables must be current).
int batch_step(fall3d_ads_ensemble *e, int iiter) {
  int rc;
  if (e->src_blocks > 0) {
    hipError_t r = source_inject_batch(e->dsrc, e->dsrc_bstart, e->nmem, e->src_blocks, e->stream);
    if (r != hipSuccess) return efail(e, FALL3D_ADS_EHIP, "batched source: %s", hipGetErrorString(r));
  }
  const int order[3] = {(iiter % 2 == 0) ? 0 : 1, (iiter % 2 == 0) ? 1 : 0, 2};
  for (int q = 0; q < 3; ++q)
    if ((rc = batch_sweep(e, order[q]))) return rc;
  hipError_t r = finalize_batch(e->djobs, e->njobs, e->stream);
  if (r != hipSuccess) return efail(e, FALL3D_ADS_EHIP, "batched reductions: %s", hipGetErrorString(r));
  return FALL3D_ADS_OK;
}

// default tile shape before any tuning
void default_tiling(fall3d_ads_ensemble *e) {
  int nxmax = 1;
  for (auto *c : e->mem) nxmax = std::max(nxmax, c->nx);
  e->x_cpt = 0;
  for (int i = 0; i < kSweepCptCount; ++i)
    if (nxmax / kSweepCpt[i] + 1 <= 64) {
      e->x_cpt = kSweepCpt[i];
      break;
    }
  // (the autotuner may also pick shapes whose widest lines take two waves:
  // their workgroups keep the 256 threads every member's launch must share)
  e->bpg[1] = e->bpg[2] = std::min(e->nb, 8);
}

// device copies of every member's mutable state, for the autotuner
struct Snapshot {
  std::vector<void *> buf;
};

int snapshot(fall3d_ads_ensemble *e, Snapshot &s, bool restore) {
  int k = 0;
  for (auto *c : e->mem) {
    void *src[4] = {c->chi, c->comp ? c->clo : nullptr, c->acum, c->acc};
    const size_t bytes[4] = {c->ncell_s * sizeof(float), c->ncell_s * sizeof(float), c->nacum * sizeof(double),
                             (size_t)ACC_LEN * sizeof(double)};
    for (int i = 0; i < 4; ++i, ++k) {
      if (!src[i]) {
        if (!restore) s.buf.push_back(nullptr);
        continue;
      }
      if (!restore) {
        void *p = nullptr;
        EHIPCHK(e, hipMalloc(&p, bytes[i]));
        s.buf.push_back(p);
        EHIPCHK(e, hipMemcpyAsync(p, src[i], bytes[i], hipMemcpyDeviceToDevice, e->stream));
      } else {
        EHIPCHK(e, hipMemcpyAsync(src[i], s.buf[k], bytes[i], hipMemcpyDeviceToDevice, e->stream));
      }
    }
  }
  EHIPCHK(e, hipStreamSynchronize(e->stream));
  return FALL3D_ADS_OK;
}

void free_snapshot(Snapshot &s) {
  for (void *p : s.buf)
    if (p) (void)hipFree(p);
  s.buf.clear();
}

// average time (ms) of `reps` launches of the batched sweep d
int time_sweep(fall3d_ads_ensemble *e, int d, int reps, double *ms) {
  hipEvent_t a, b;
  EHIPCHK(e, hipEventCreate(&a));
  EHIPCHK(e, hipEventCreate(&b));
  int rc = batch_sweep(e, d);  // warm-up
  if (!rc) {
    (void)hipEventRecord(a, e->stream);
    for (int r = 0; r < reps && !rc; ++r) rc = batch_sweep(e, d);
    (void)hipEventRecord(b, e->stream);
    (void)hipEventSynchronize(b);
    float t = 0.0f;
    (void)hipEventElapsedTime(&t, a, b);
    *ms = t / reps;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return rc;
}

}  // namespace

extern "C" {

int fall3d_ads_ensemble_create(fall3d_ads_ensemble **pens, int nmembers, const int *dims, int nbins, int limiter,
                               int time_marching, int real_kind, int state_mode) {
  if (!pens) return FALL3D_ADS_EINVAL;
  *pens = nullptr;
  if (nmembers < 1 || !dims) return FALL3D_ADS_EINVAL;
  if (state_mode == FALL3D_ADS_STATE_FP64) return FALL3D_ADS_EINVAL;  // members run the FP32 states
  auto *e = new fall3d_ads_ensemble();
  e->nmem = nmembers;
  e->nb = nbins;
  e->limiter = limiter;
  e->tm = time_marching;
  e->kind = real_kind;
  e->comp = state_mode == FALL3D_ADS_STATE_FP32X2;
  *pens = e;
  if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess)
    return efail(e, FALL3D_ADS_EHIP, "stream creation failed");
  e->own_stream = true;
  for (int m = 0; m < nmembers; ++m) {
    fall3d_ads_ctx *c = nullptr;
    int rc = fall3d_ads_initialize(&c, dims[3 * m], dims[3 * m + 1], dims[3 * m + 2], nbins, limiter,
                                   time_marching, real_kind, state_mode);
    if (rc) {
      const std::string msg = c ? fall3d_ads_last_error(c) : "initialisation failed";
      if (c) fall3d_ads_release(c);
      return efail(e, rc, "member %d: %s", m, msg.c_str());
    }
    e->mem.push_back(c);
    if ((rc = fall3d_ads_set_stream(c, e->stream))) return mfail(e, m, rc);
  }
  for (int d = 0; d < 3; ++d) {
    EHIPCHK(e, dalloc(&e->dargs[d], nmembers));
    EHIPCHK(e, dalloc(&e->dbstart[d], nmembers + 1));
  }
  EHIPCHK(e, dalloc(&e->dsrc, nmembers));
  EHIPCHK(e, dalloc(&e->dsrc_bstart, nmembers + 1));
  default_tiling(e);
  if (e->x_cpt == 0) return efail(e, FALL3D_ADS_EINVAL, "no supported x tiling (nx <= 1024 required)");
  return apply_tiling(e);
}

int fall3d_ads_ensemble_release(fall3d_ads_ensemble *e) {
  if (!e) return FALL3D_ADS_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  drop_graph(e);
  for (auto *c : e->mem) {
    // the members borrow the ensemble stream: detach before it is destroyed
    c->stream = nullptr;
    c->own_stream = false;
    fall3d_ads_release(c);
  }
  for (int d = 0; d < 3; ++d) {
    dfree(e->dargs[d]);
    dfree(e->dbstart[d]);
  }
  dfree(e->dsrc);
  dfree(e->dsrc_bstart);
  dfree(e->djobs);
  if (e->own_stream && e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return FALL3D_ADS_OK;
}

const char *fall3d_ads_ensemble_last_error(const fall3d_ads_ensemble *e) {
  return e ? e->err.c_str() : "null ensemble";
}

int fall3d_ads_ensemble_member(fall3d_ads_ensemble *e, int m, fall3d_ads_ctx **ctx) {
  if (!e || !ctx || m < 0 || m >= e->nmem) return FALL3D_ADS_EINVAL;
  *ctx = e->mem[m];
  return FALL3D_ADS_OK;
}

int fall3d_ads_ensemble_set_tiling(fall3d_ads_ensemble *e, int x_cells_per_lane, int y_bins_per_group,
                                   int z_bins_per_group) {
  if (!e) return FALL3D_ADS_EINVAL;
  const int ox = e->x_cpt, oy = e->bpg[1], oz = e->bpg[2];
  if (x_cells_per_lane) {
    if (!sweep_cpt_supported(x_cells_per_lane))
      return efail(e, FALL3D_ADS_EINVAL, "unsupported cells per lane %d", x_cells_per_lane);
    for (auto *c : e->mem)
      if (c->nx / x_cells_per_lane + 1 > 128)  // at most two waves per line: 256-thread workgroups
        return efail(e, FALL3D_ADS_EINVAL, "cells per lane %d too small for nx = %d", x_cells_per_lane, c->nx);
    e->x_cpt = x_cells_per_lane;
  }
  if (y_bins_per_group) e->bpg[1] = std::max(1, std::min(y_bins_per_group, 8));
  if (z_bins_per_group) e->bpg[2] = std::max(1, std::min(z_bins_per_group, 8));
  const int rc = apply_tiling(e);
  if (rc) {
    e->x_cpt = ox;
    e->bpg[1] = oy;
    e->bpg[2] = oz;
    (void)apply_tiling(e);
  }
  return rc;
}

int fall3d_ads_ensemble_get_tiling(fall3d_ads_ensemble *e, int out[3]) {
  if (!e || !out) return FALL3D_ADS_EINVAL;
  out[0] = e->x_cpt;
  out[1] = e->bpg[1];
  out[2] = e->bpg[2];
  return FALL3D_ADS_OK;
}

int fall3d_ads_ensemble_step(fall3d_ads_ensemble *e, const double *dt, const double *stime, int iiter) {
  if (!e || !dt) return FALL3D_ADS_EINVAL;
  int rc;
  if ((rc = members_ready(e))) return rc;
  std::vector<double> s0(e->nmem, 0.0);
  const double *st = stime ? stime : s0.data();
  if ((rc = build_tables(e, dt, st))) return rc;
  return batch_step(e, iiter);
}

int fall3d_ads_ensemble_run(fall3d_ads_ensemble *e, const double *dt, const double *stime, int iiter0,
                            int nsteps) {
  if (!e || !dt || nsteps < 0) return FALL3D_ADS_EINVAL;
  int rc;
  if ((rc = members_ready(e))) return rc;
  std::vector<double> s0(e->nmem, 0.0);
  const double *st = stime ? stime : s0.data();
  if ((rc = build_tables(e, dt, st))) return rc;
  if (!e->use_graph || nsteps < 2) {
    for (int s = 0; s < nsteps; ++s)
      if ((rc = batch_step(e, iiter0 + s))) return rc;
    return FALL3D_ADS_OK;
  }
  // the whole run as one graph (re-captured when the step count, the
  // parity of the first iteration, a time step or a member changes)
  bool same = e->gexec && e->g_steps == nsteps && e->g_iiter_parity == (iiter0 & 1);
  for (int m = 0; same && m < e->nmem; ++m)
    same = e->g_dt[m] == dt[m] && e->g_stime[m] == st[m] && e->g_version[m] == e->mem[m]->version;
  if (!same) {
    drop_graph(e);
    EHIPCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int r = FALL3D_ADS_OK;
    for (int s = 0; s < nsteps && !r; ++s) r = batch_step(e, iiter0 + s);
    hipGraph_t gr = nullptr;
    hipError_t ce = hipStreamEndCapture(e->stream, &gr);
    if (r) {
      if (gr) (void)hipGraphDestroy(gr);
      return r;
    }
    if (ce != hipSuccess) return efail(e, FALL3D_ADS_EHIP, "ensemble capture: %s", hipGetErrorString(ce));
    ce = hipGraphInstantiate(&e->gexec, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (ce != hipSuccess) {
      e->gexec = nullptr;
      return efail(e, FALL3D_ADS_EHIP, "ensemble graph instantiate: %s", hipGetErrorString(ce));
    }
    e->g_steps = nsteps;
    e->g_iiter_parity = iiter0 & 1;
    e->g_dt.assign(dt, dt + e->nmem);
    e->g_stime.assign(st, st + e->nmem);
    e->g_version.resize(e->nmem);
    for (int m = 0; m < e->nmem; ++m) e->g_version[m] = e->mem[m]->version;
  }
  EHIPCHK(e, hipGraphLaunch(e->gexec, e->stream));
  return FALL3D_ADS_OK;
}

int fall3d_ads_ensemble_set_graph(fall3d_ads_ensemble *e, int on) {
  if (!e) return FALL3D_ADS_EINVAL;
  e->use_graph = on != 0;
  if (!e->use_graph) drop_graph(e);
  return FALL3D_ADS_OK;
}

int fall3d_ads_ensemble_synchronize(fall3d_ads_ensemble *e) {
  if (!e) return FALL3D_ADS_EINVAL;
  EHIPCHK(e, hipStreamSynchronize(e->stream));
  return FALL3D_ADS_OK;
}

int fall3d_ads_ensemble_autotune(fall3d_ads_ensemble *e, const double *dt, const double *stime, int reps,
                                 double *best_ms) {
  if (!e || !dt) return FALL3D_ADS_EINVAL;
  int rc;
  if ((rc = members_ready(e))) return rc;
  if (reps < 1) reps = 3;
  std::vector<double> s0(e->nmem, 0.0);
  const double *st = stime ? stime : s0.data();
  Snapshot snap;
  if ((rc = snapshot(e, snap, false))) {
    free_snapshot(snap);
    return rc;
  }
  // candidates: x cells per lane (every supported value that fits the
  // widest member); y/z bins per group g in 1..min(nb, 8) and bins per
  // thread b (2 not with the compensated state), encoded g + 16 (b - 1)
  int nxmax = 1;
  bool comp = false;
  for (auto *c : e->mem) {
    nxmax = std::max(nxmax, c->nx);
    comp = comp || c->comp;
  }
  std::vector<int> xc, yz;
  for (int i = 0; i < kSweepCptCount; ++i)
    if (nxmax / kSweepCpt[i] + 1 <= 128) xc.push_back(kSweepCpt[i]);  // up to two waves per line
  for (int b = 1; b <= (comp ? 1 : 2); ++b)
    for (int g = b; g <= std::min(e->nb, 8); ++g) {
      const int groups = (e->nb + g - 1) / g;
      if ((e->nb + groups - 1) / groups == g) yz.push_back(g + 16 * (b - 1));  // distinct shapes only
    }
  double best[3] = {1e30, 1e30, 1e30};
  int pick[3] = {e->x_cpt, e->bpg[1] + 16 * std::max(e->bpt[1] - 1, 0), e->bpg[2] + 16 * std::max(e->bpt[2] - 1, 0)};
  for (int d = 0; d < 3 && !rc; ++d) {
    const std::vector<int> &cand = (d == 0) ? xc : yz;
    for (int v : cand) {
      if (d == 0) {
        e->x_cpt = v;
      } else {
        e->bpg[d] = v & 15;
        e->bpt[d] = (v >> 4) + 1;
      }
      if ((rc = apply_tiling(e))) break;
      if ((rc = build_tables(e, dt, st))) break;
      double ms = 0.0;
      if ((rc = time_sweep(e, d, reps, &ms))) break;
      if (ms < best[d]) {
        best[d] = ms;
        pick[d] = v;
      }
    }
    if (d == 0) {
      e->x_cpt = pick[0];
    } else {
      e->bpg[d] = pick[d] & 15;
      e->bpt[d] = (pick[d] >> 4) + 1;
    }
  }
  const int rr = snapshot(e, snap, true);
  free_snapshot(snap);
  if (rc) return rc;
  if (rr) return rr;
  if ((rc = apply_tiling(e))) return rc;
  if (best_ms)
    for (int d = 0; d < 3; ++d) best_ms[d] = best[d];
  return FALL3D_ADS_OK;
}

}  // extern "C"
