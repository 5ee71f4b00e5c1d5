// ads_context.h -- the context object behind the C ABI (fall3d_ads_ctx) and
// the host helpers that the single-domain and ensemble entry points share.
// Internal header of libfall3d_ads_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/fall3d_ads.h"
#include "ads_aux.h"
#include "ads_internal.h"

namespace f3d {
namespace host {

enum { KID_SOURCE = 0, KID_SWEEP_X = 1, KID_SWEEP_Y = 2, KID_SWEEP_Z = 3, KID_REDUCE = 4, KID_COUNT = 5 };

// device accumulator layout
// ACC_RESIDUAL + d: FP32 state only, the mass the stored FP32 values lost
// (or gained) to rounding in the final updates of sweep d, sum of
// (exact - stored) dxb area. Every reduction job of a step owns its
// destination slots: the jobs run concurrently in one launch.
enum { ACC_FLUX = 0, ACC_LEDGER = 6, ACC_INJECTED = 12, ACC_RESIDUAL = 13, ACC_LEN = 16 };

struct Tiling {
  int cpt = 0;      // cells per lane (x sweep) / bins per thread (y, z sweeps)
  int S = 0;        // segments per line
  int wpl = 1;      // x sweep: waves per line
  int TX = 0;       // lines per workgroup (y, z) / lines per workgroup (x)
  int block = 0;    // threads per workgroup
  int gx = 0;       // workgroups along grid.x
  int groups = 1;   // bin groups (grid.y)
  int bpg = 1;      // bins per group
  int ntrans = 0;   // x tiles (y, z sweeps)
};

struct PendingEvent {
  int kid;
  hipEvent_t a, b;
};

}  // namespace host
}  // namespace f3d

struct fall3d_ads_ctx {
  int nx = 0, ny = 0, nz = 0, nb = 0;
  int limiter = 1, tm = 2, kind = 8, comp = 0;
  int bc = 0;  // boundary rule of the sweeps (FALL3D_ADS_BC_*)
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string err = "no error";

  // state
  // c (and its low words) start kStateLead floats into their allocations,
  // with rows of state_pitch(nx) floats: cell 1 of every row starts a
  // 128-byte line (the sweeps read and write rows in 16-byte pieces)
  float *chi = nullptr, *clo = nullptr;
  float *chi_base = nullptr, *clo_base = nullptr;
  // FP64 state (FALL3D_ADS_STATE_FP64): c64 holds the concentration in place
  // of chi/clo, and the sweeps read FP64 copies of the fields below in the
  // same device layouts (the FP32 copies remain for the CFL time step)
  double *c64 = nullptr;
  double *u1d = nullptr, *u2d = nullptr, *v1d = nullptr, *v2d = nullptr, *w1d = nullptr, *w2d = nullptr;
  double *k1d = nullptr, *k2d = nullptr, *k3d = nullptr, *vsd = nullptr;
  // the FP64 x sweep runs on (x, y)-transposed copies of the state and of its
  // fields (y fastest), so that the lanes of a wave (consecutive y lines)
  // read adjacent words: c64t (-1:ny+2,-1:nx+2,-1:nz+2,nb), u1t/u2t
  // (ny,0:nx+2,nz), k1t (ny,-1:nx+2,nz)
  double *c64t = nullptr, *u1t = nullptr, *u2t = nullptr, *k1t = nullptr;
  double *acum = nullptr;
  size_t ncell_c = 0;  // (nx+4)(ny+4)(nz+4)nb: cells of the state with ghosts (the host layout, c64)
  size_t ncell_s = 0;  // state_pitch(nx)(ny+4)(nz+4)nb: words of chi and of clo
  size_t nacum = 0;    // (nx+4)(ny+4)nb

  // meteo (device layouts, FP32)
  float *u1 = nullptr, *u2 = nullptr, *v1 = nullptr, *v2 = nullptr, *w1 = nullptr, *w2 = nullptr;
  float *k1 = nullptr, *k2 = nullptr, *k3 = nullptr, *vs = nullptr;
  bool meteo_ok = false;
  bool has_w2 = false;

  // grid
  double *dxp64[3] = {nullptr, nullptr, nullptr};
  double *dxb64[3] = {nullptr, nullptr, nullptr};
  float *hdxp[3] = {nullptr, nullptr, nullptr};
  float *idxb[3] = {nullptr, nullptr, nullptr};
  float *dttab[3] = {nullptr, nullptr, nullptr};  // dt/dxb, 0.5 dt/dxb, final factor (for dtib_dt)
  float *itab[3] = {nullptr, nullptr, nullptr};   // per-iteration factors of the streaming sweeps
  double *dttab64[3] = {nullptr, nullptr, nullptr};  // final-update factor w dt/dxb in FP64 (for dtib_dt)
  double dtib_dt = 0.0;
  bool dtib_ok = false;
  std::vector<double> hdxp64[3], hdxb64[3];
  bool grid_ok = false;

  // sparse source
  int nsrc = 0;
  size_t *src_idx = nullptr;
  double *src_rate = nullptr, *src_meas = nullptr;
  double *src_part = nullptr;
  double *src_s = nullptr, *src_vol = nullptr;  // FP64 state: rate and cell volume of each entry

  // accumulators and reduction scratch
  double *acc = nullptr;
  double *part[4] = {nullptr, nullptr, nullptr, nullptr};  // source, x, y, z
  size_t part_cap[4] = {0, 0, 0, 0};                        // in blocks
  double *rpart[4] = {nullptr, nullptr, nullptr, nullptr};  // rounding-residual partials of x, y, z (1..3)
  double *red_part = nullptr;
  int red_cap = 0;

  // staging
  void *dstage = nullptr;
  size_t dstage_bytes = 0;
  void *hu = nullptr, *hv = nullptr, *hw = nullptr;  // CB%u, CB%v, CB%w host arrays

  // tiling
  f3d::host::Tiling tl[3];
  int user_cpt[3] = {0, 0, 0}, user_tx[3] = {0, 0, 0}, user_bpg[3] = {0, 0, 0};

  // graph replay of step pairs (even + odd iteration)
  bool use_graph = true;
  hipGraphExec_t gexec = nullptr;
  double g_dt = -1.0, g_stime = -1.0;
  int g_version = -1;
  int version = 0;  // bumped whenever launch parameters of a step change

  // profiling
  bool prof = false;    // the step being issued records per-kernel events
  int prof_every = 0;   // fall3d_ads_run times one step in prof_every (0: none)
  std::vector<f3d::host::PendingEvent> pending;
  std::vector<hipEvent_t> evpool;
  double kt_ms[f3d::host::KID_COUNT] = {0, 0, 0, 0, 0};
  long long kt_n[f3d::host::KID_COUNT] = {0, 0, 0, 0, 0};

  // optional outputs/extensions of the x and y sweeps (ads_learned.hip)
  float *fout[2] = {nullptr, nullptr};  // face mass flux per unit area, accumulated (u / v layout per bin)
  bool fout_on = false;
  struct fall3d_mlp *learned = nullptr;  // flux-correction network (borrowed) or nullptr: plain KT
  float *dphi[2] = {nullptr, nullptr};   // its face corrections of the current step
  int *lact[2] = {nullptr, nullptr};     // active 16-face tiles of each axis (compacted list)
  unsigned char *lflag[2] = {nullptr, nullptr};  // per tile: 1 = active this step
  unsigned long long *lbits = nullptr;  // per cell row, bit i-1 of word (i-1)/64: |c(i)| above the threshold
  int *lseg = nullptr;                   // 64-cell row segments next to an active tile (compacted list)
  unsigned *lcount = nullptr;            // counts of the three lists
  unsigned *lbad = nullptr;              // [0] non-finite network output this step, [1] fallback steps
  double learned_smin = 0.0;
  double learned_gain = 1.0;
  double learned_flux_limit = 0.0;       // |dPhi| <= limit * |plain face flux| (0: off)
  float *lthr = nullptr;                 // device threshold (relative smin) and its partials
  int lthr_blocks = 0;
};

namespace f3d {
namespace host {

int fail(fall3d_ads_ctx *c, int code, const char *fmt, ...);

#define HIPCHK(ctx, expr)                                                                              \
  do {                                                                                                 \
    hipError_t e_ = (expr);                                                                            \
    if (e_ != hipSuccess) return f3d::host::fail((ctx), FALL3D_ADS_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

template <typename T>
hipError_t dalloc(T **p, size_t n) {
  // kDallocPad zeroed elements after the end: the x sweep reads a lane's
  // cells with wide loads that may run up to 17 elements past a line's last
  // face, which for the last line of an array is past its end
  constexpr size_t kDallocPad = 64;
  *p = nullptr;
  hipError_t e = hipMalloc((void **)p, (n + kDallocPad) * sizeof(T));
  if (e == hipSuccess) e = hipMemset((void *)(*p + n), 0, kDallocPad * sizeof(T));
  return e;
}

inline void dfree(void *p) {
  if (p) (void)hipFree(p);
}

inline int axis_n(const fall3d_ads_ctx *c, int dir) { return dir == 0 ? c->nx : (dir == 1 ? c->ny : c->nz); }
inline bool is_fp64(const fall3d_ads_ctx *c) { return c->comp == FALL3D_ADS_STATE_FP64; }
// lines of a sweep along dir (per bin)
inline long axis_lines(const fall3d_ads_ctx *c, int dir) {
  return dir == 0 ? (long)c->ny * c->nz : (dir == 1 ? (long)c->nx * c->nz : (long)c->nx * c->ny);
}

// dt/dxb tables of the three axes for time step dt (stream-ordered)
int ensure_dt(fall3d_ads_ctx *c, double dt);
// kernel arguments of one sweep of context c under its current tiling, with
// the context's resident winds and diffusion (or the given ones)
void make_sweep_args(fall3d_ads_ctx *c, int dir, double dt, double stime, int write_ghosts, SweepArgs &a);
void make_sweep_args(fall3d_ads_ctx *c, int dir, double dt, double stime, int write_ghosts, const float *vel1,
                     const float *vel2, const float *kd, SweepArgs &a);
// tiling of one axis for context c; see choose_tiling in ads_capi.hip
Tiling choose_tiling(const fall3d_ads_ctx *c, int dir);
int refresh_tiling(fall3d_ads_ctx *c);

// learned flux correction around one step (ads_learned.hip): corrections
// from the state at the start of the step, applied after the plain step
int learned_pre_step(fall3d_ads_ctx *c, double dt, double stime);
int learned_post_step(fall3d_ads_ctx *c);
void learned_release(fall3d_ads_ctx *c);

}  // namespace host
}  // namespace f3d
