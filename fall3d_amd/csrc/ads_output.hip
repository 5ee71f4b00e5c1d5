// ads_output.hip -- the output post-processing of nc_IO_out_res
// (Sources/mod_nc_IO.F90:3206-4335) on the device, for the context's state:
//
//   c scaled back by the map factors, converted to g/m3 and clipped at 0 on
//   the mass points with their 2 ghosts (map factors at the clamped index)
//                                                        mod_nc_IO.F90:3397-3416
//   mass points -> cell corners, 8-point average          mod_Grid.F90:1150-1181
//   ground accumulation -> corners, 4-point average of the clamped field
//                                                        mod_Grid.F90:1190-1220
//   per species: 3-D concentration (CON), surface concentration (CON_SUR),
//   column load by the trapezoid rule over the corner heights (COL_MASS,
//   optionally in Dobson units), PM column loads (COL_MASS_PM), cloud-top
//   height at 50 ug/m3 (CLOUD_TOP), ground load total and per bin (GRN_LOAD,
//   GRN_LOAD_BIN)                                        mod_nc_IO.F90:3506-4020
//
// One thread owns one corner column (i, j) and walks it bottom to top, so
// every sum runs in the reference's order (bins in order, levels upward);
// neighbouring threads own neighbouring columns (coalesced along x).
#include <cmath>
#include <cstring>
#include <vector>

#include "ads_aux.h"
#include "ads_context.h"

namespace f3d {
namespace {

struct OutArgs {
  const float *chi, *clo;  // FP32 state (lo: compensated) or
  const double *c64;       // FP64 state
  const double *acum;      // acum(-1:nx+2,-1:ny+2,nb)
  int nx, ny, nz, nb;
  const double *hm1, *hm2, *hm3;  // Hm1_p(1:ny), Hm2_p(-1:ny+2), Hm3_p(1:nx,1:ny)
  const double *zc;               // corner heights (nx+1)(ny+1)(nz+1)
  const int *spe;                 // per bin species code
  const int *pm_bin;              // per bin: bit p set when the bin counts for PM threshold p
  int species, so2_du, npm, pm_ok;
  // outputs (corner layouts, FP64; any may be nullptr)
  double *con_bins;     // (nx+1)(ny+1)(nz+1) nb
  double *con;          // (nx+1)(ny+1)(nz+1)
  double *con_sur;      // (nx+1)(ny+1)
  double *col_mass;     // (nx+1)(ny+1)
  double *col_mass_pm;  // (nx+1)(ny+1) npm
  double *cloud_top;    // (nx+1)(ny+1)
  double *grn_load;     // (nx+1)(ny+1)
  double *grn_load_bin; // (nx+1)(ny+1) nb
};

constexpr int kMaxPm = 8;

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// scaled-back, clipped concentration in g/m3 at mass point (i, j, k) of bin b,
// i in 0..nx+1, j in 0..ny+1, k in 0..nz+1
__device__ __forceinline__ double cs_at(const OutArgs &a, int i, int j, int k, int b) {
  const int px = a.c64 ? a.nx + 4 : state_pitch(a.nx);  // x pitch of c64 / chi
  const size_t ci = (size_t)(i + 1) + (size_t)px * ((size_t)(j + 1) + (size_t)(a.ny + 4) *
                                                    ((size_t)(k + 1) + (size_t)(a.nz + 4) * b));
  double c;
  if (a.c64) {
    c = a.c64[ci];
  } else {
    c = (double)a.chi[ci];
    if (a.clo) c += (double)a.clo[ci];
  }
  const int jj = clampi(j, 1, a.ny), ii = clampi(i, 1, a.nx);
  const double hm1 = a.hm1[jj - 1], hm2 = a.hm2[jj + 1], hm3 = a.hm3[(size_t)(ii - 1) + (size_t)a.nx * (jj - 1)];
  const double v = 1e3 * c / (hm1 * hm2 * hm3);
  return fmax(0.0, v);
}

__device__ __forceinline__ double cc_at(const OutArgs &a, int i, int j, int k, int b) {
  return (cs_at(a, i, j, k, b) + cs_at(a, i - 1, j, k, b) + cs_at(a, i, j - 1, k, b) + cs_at(a, i - 1, j - 1, k, b) +
          cs_at(a, i, j, k - 1, b) + cs_at(a, i - 1, j, k - 1, b) + cs_at(a, i, j - 1, k - 1, b) +
          cs_at(a, i - 1, j - 1, k - 1, b)) /
         8.0;
}

__global__ void __launch_bounds__(256) k_output(OutArgs a) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t ncol = (size_t)(nx + 1) * (ny + 1);
  const size_t n3 = ncol * (nz + 1);
  for (size_t col = blockIdx.x * (size_t)blockDim.x + threadIdx.x; col < ncol; col += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(col % (nx + 1)) + 1;
    const int j = (int)(col / (nx + 1)) + 1;
    auto zc = [&](int k) { return a.zc ? a.zc[col + ncol * (size_t)(k - 1)] : 0.0; };
    double load = 0.0, top = 0.0, wprev = 0.0, zprev = 0.0;
    double pm[kMaxPm], pmprev[kMaxPm];
    for (int p = 0; p < kMaxPm; ++p) pm[p] = pmprev[p] = 0.0;
    for (int k = 1; k <= nz + 1; ++k) {
      double w = 0.0, wp[kMaxPm];
      for (int p = 0; p < kMaxPm; ++p) wp[p] = 0.0;
      for (int b = 0; b < a.nb; ++b) {
        const double v = cc_at(a, i, j, k, b);
        if (a.con_bins) a.con_bins[col + ncol * (size_t)(k - 1) + n3 * (size_t)b] = v;
        if (a.spe[b] != a.species) continue;
        w = w + v;
        if (a.pm_ok)
          for (int p = 0; p < a.npm; ++p)
            if ((a.pm_bin[b] >> p) & 1) wp[p] = wp[p] + v;
      }
      if (a.con) a.con[col + ncol * (size_t)(k - 1)] = w;
      if (k == 1 && a.con_sur) a.con_sur[col] = w;
      const double z = zc(k);
      if (k > 1) {
        load = load + 0.5 * (wprev + w) * (z - zprev);
        if (a.pm_ok)
          for (int p = 0; p < a.npm; ++p) pm[p] = pm[p] + 0.5 * (pmprev[p] + wp[p]) * (z - zprev);
      }
      if (w > 50e-6) top = z;  // the highest level above the threshold wins
      wprev = w;
      zprev = z;
      for (int p = 0; p < kMaxPm; ++p) pmprev[p] = wp[p];
    }
    if (a.so2_du) load = load / 64.0 * 2.238e3;
    if (a.col_mass) a.col_mass[col] = load;
    if (a.cloud_top) a.cloud_top[col] = top;
    if (a.col_mass_pm && a.pm_ok)
      for (int p = 0; p < a.npm; ++p) a.col_mass_pm[col + ncol * (size_t)p] = pm[p];
    // ground accumulation at the corner (clamped mass points), kg/m2
    double g = 0.0;
    const int i0 = clampi(i - 1, 1, nx), i1 = clampi(i, 1, nx), j0 = clampi(j - 1, 1, ny), j1 = clampi(j, 1, ny);
    for (int b = 0; b < a.nb; ++b) {
      auto A = [&](int ii, int jj) {
        return a.acum[(size_t)(ii + 1) + (size_t)(nx + 4) * ((size_t)(jj + 1) + (size_t)(ny + 4) * b)];
      };
      const double ac = (A(i1, j1) + A(i0, j1) + A(i1, j0) + A(i0, j0)) / 4.0;
      if (a.grn_load_bin) a.grn_load_bin[col + ncol * (size_t)b] = ac;
      if (a.spe[b] == a.species) g = g + ac;
    }
    if (a.grn_load) a.grn_load[col] = g;
  }
}

// grid_get_shapez (Sources/mod_Grid.F90:1287-1329) over v(1..n) (element
// q at v[(q-1) * stride]): the interval [v(k), v(k+1)] holding z, searched
// from the bottom, and the linear weight of v(k+1); below v(1) it gives
// (1, 0), above v(n) (n-1, 1).
__host__ __device__ inline void shapez(const double *v, size_t stride, int n, double z, int &ik, double &s) {
  if (z < v[0]) {
    ik = 1;
    s = 0.0;
    return;
  }
  if (z > v[(size_t)(n - 1) * stride]) {
    ik = n - 1;
    s = 1.0;
    return;
  }
  for (int k = 1; k < n; ++k) {
    const double lo = v[(size_t)(k - 1) * stride], hi = v[(size_t)k * stride];
    if (z >= lo && z <= hi) {
      ik = k;
      s = (z - lo) / (hi - lo);
      return;
    }
  }
  ik = 1;  // only for a column that is not ordered upward (the reference never ends its search then)
  s = 0.0;
}

// the species' corner concentration on the z levels zl(1..nz+1) above sea
// level: per corner column, interpolated over the column's own corner
// heights (mod_nc_IO.F90:3302-3313 and 4102-4117)
__global__ void __launch_bounds__(256) k_zlevels(const double *__restrict__ con, const double *__restrict__ zc,
                                                 const double *__restrict__ zl, int nx, int ny, int nz,
                                                 double *__restrict__ cz) {
  const size_t ncol = (size_t)(nx + 1) * (ny + 1);
  for (size_t col = blockIdx.x * (size_t)blockDim.x + threadIdx.x; col < ncol; col += (size_t)gridDim.x * blockDim.x) {
    for (int k = 1; k <= nz + 1; ++k) {
      int kk;
      double sh;
      shapez(zc + col, ncol, nz + 1, zl[k - 1], kk, sh);
      cz[col + ncol * (size_t)(k - 1)] =
          kk > 0 ? (1.0 - sh) * con[col + ncol * (size_t)(kk - 1)] + sh * con[col + ncol * (size_t)kk] : 0.0;
    }
  }
}

// Cuts of cz (mod_nc_IO.F90:4126-4332): mode 0 x cuts (ny+1, nz+1, ncut),
// 1 y cuts (nx+1, nz+1, ncut), 2 z or flight-level cuts (nx+1, ny+1, ncut);
// cut q interpolates between positions idx[q] and idx[q]+1 with weight sh[q].
__global__ void __launch_bounds__(256) k_cuts(const double *__restrict__ cz, int nx, int ny, int nz, int mode,
                                              int ncut, const int *__restrict__ idx, const double *__restrict__ sh,
                                              double *__restrict__ out) {
  const size_t n1 = (size_t)(mode == 0 ? ny + 1 : nx + 1), n2 = (size_t)(mode == 2 ? ny + 1 : nz + 1);
  const size_t total = n1 * n2 * (size_t)ncut;
  const size_t sx = 1, sy = (size_t)(nx + 1), sz = (size_t)(nx + 1) * (ny + 1);
  for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
    const size_t a = e % n1, r = e / n1, b = r % n2;
    const int q = (int)(r / n2);
    const size_t m = (size_t)(idx[q] - 1);
    size_t lo, step;
    if (mode == 0) {  // (j, k) at x position m
      lo = m * sx + a * sy + b * sz;
      step = sx;
    } else if (mode == 1) {  // (i, k) at y position m
      lo = a * sx + m * sy + b * sz;
      step = sy;
    } else {  // (i, j) at level m
      lo = a * sx + b * sy + m * sz;
      step = sz;
    }
    out[e] = (1.0 - sh[q]) * cz[lo] + sh[q] * cz[lo + step];
  }
}

}  // namespace
}  // namespace f3d

using namespace f3d;
using namespace f3d::host;

namespace {

// cut request of fall3d_ads_postprocess_cuts: host outputs x, y, z, FL
struct CutReq {
  const fall3d_cut_args *args;
  double *out[4];
};

int postprocess_impl(fall3d_ads_ctx *c, const fall3d_output_args *o, double *con_bins, double *con, double *con_sur,
                     double *col_mass, double *col_mass_pm, double *cloud_top, double *grn_load,
                     double *grn_load_bin, const CutReq *cut) {
  if (!c) return FALL3D_ADS_EINVAL;
  if (!o) return fail(c, FALL3D_ADS_EINVAL, "null argument");
  if (cut) {
    const fall3d_cut_args *ca = cut->args;
    if (!ca) return fail(c, FALL3D_ADS_EINVAL, "null cut arguments");
    if (!o->z_c || !ca->sigma) return fail(c, FALL3D_ADS_EINVAL, "cuts need the corner heights z_c and sigma levels");
    const int ncut[4] = {ca->ncutx, ca->ncuty, ca->ncutz, ca->nfl};
    const double *pos[4] = {ca->x_cut, ca->y_cut, ca->z_cut, ca->fl_cut};
    for (int q = 0; q < 4; ++q)
      if (ncut[q] < 0 || (ncut[q] > 0 && (!pos[q] || !cut->out[q])))
        return fail(c, FALL3D_ADS_EINVAL, "cut kind %d: positive count needs positions and an output", q);
  }
  if (!c->grid_ok) return fail(c, FALL3D_ADS_ESTATE, "fall3d_ads_set_grid has not been called");
  if (o->npm < 0 || o->npm > kMaxPm) return fail(c, FALL3D_ADS_EINVAL, "at most %d PM thresholds", kMaxPm);
  if ((col_mass || col_mass_pm || cloud_top) && !o->z_c)
    return fail(c, FALL3D_ADS_EINVAL, "column loads and cloud top need the corner heights z_c");
  if (col_mass_pm && o->npm > 0 && (!o->pm_um || !o->bin_diam))
    return fail(c, FALL3D_ADS_EINVAL, "PM loads need the thresholds and the bin diameters");
  const int nx = c->nx, ny = c->ny, nz = c->nz, nb = c->nb;
  const size_t ncol = (size_t)(nx + 1) * (ny + 1), n3 = ncol * (nz + 1);
  // host-side tables: map factors (1 when not given), species, PM membership
  std::vector<double> hm1(ny, 1.0), hm2(ny + 4, 1.0), hm3((size_t)nx * ny, 1.0), zc;
  auto hv = [&](const void *p, size_t q) {
    return c->kind == 8 ? ((const double *)p)[q] : (double)((const float *)p)[q];
  };
  if (o->Hm1_p) for (int q = 0; q < ny; ++q) hm1[q] = hv(o->Hm1_p, q);
  if (o->Hm2_p) for (int q = 0; q < ny + 4; ++q) hm2[q] = hv(o->Hm2_p, q);
  if (o->Hm3_p) for (size_t q = 0; q < hm3.size(); ++q) hm3[q] = hv(o->Hm3_p, q);
  if (o->z_c) {
    zc.resize(n3);
    for (size_t q = 0; q < n3; ++q) zc[q] = hv(o->z_c, q);
  }
  std::vector<int> spe(nb, 0), pmb(nb, 0);
  for (int b = 0; b < nb; ++b) spe[b] = o->bin_species ? o->bin_species[b] : 0;
  const int species = o->bin_species ? o->species : 0;
  // PM loads skip aerosol species (mod_nc_IO.F90:3750); a bin counts for
  // threshold p when its diameter is at most pm_um[p] micrometres
  const bool pm_ok = col_mass_pm && o->npm > 0 && o->species_category != 2;
  if (pm_ok)
    for (int b = 0; b < nb; ++b)
      for (int p = 0; p < o->npm; ++p)
        if (o->bin_diam[b] <= o->pm_um[p] * 1e-6) pmb[b] |= 1 << p;

  // device buffers: tables and outputs
  std::vector<void *> owned;
  auto cleanup = [&]() {
    for (void *p : owned) dfree(p);
  };
  auto dev = [&](size_t bytes) -> void * {
    char *p = nullptr;
    if (dalloc(&p, bytes) != hipSuccess) return nullptr;
    owned.push_back(p);
    return p;
  };
  OutArgs a;
  memset(&a, 0, sizeof(a));
  a.chi = c->chi;
  a.clo = c->comp == FALL3D_ADS_STATE_FP32X2 ? c->clo : nullptr;
  a.c64 = c->c64;
  a.acum = c->acum;
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  a.nb = nb;
  a.species = species;
  a.so2_du = o->so2_du;
  a.npm = o->npm;
  a.pm_ok = pm_ok ? 1 : 0;
  double *d_hm1 = (double *)dev(8 * hm1.size()), *d_hm2 = (double *)dev(8 * hm2.size());
  double *d_hm3 = (double *)dev(8 * hm3.size());
  int *d_spe = (int *)dev(4 * (size_t)nb), *d_pmb = (int *)dev(4 * (size_t)nb);
  double *d_zc = o->z_c ? (double *)dev(8 * n3) : nullptr;
  struct Out {
    double *host;
    size_t n;
    double *dev;
  } outs[8] = {{con_bins, n3 * nb, nullptr}, {con, n3, nullptr},           {con_sur, ncol, nullptr},
               {col_mass, ncol, nullptr},    {col_mass_pm, ncol * (size_t)std::max(o->npm, 1), nullptr},
               {cloud_top, ncol, nullptr},   {grn_load, ncol, nullptr},     {grn_load_bin, ncol * nb, nullptr}};
  bool ok = d_hm1 && d_hm2 && d_hm3 && d_spe && d_pmb && (!o->z_c || d_zc);
  for (auto &x : outs)
    if ((x.host || (cut && &x == &outs[1])) && ok) ok = (x.dev = (double *)dev(8 * x.n)) != nullptr;
  double *d_cz = nullptr, *d_zl = nullptr;
  if (cut && ok) ok = (d_cz = (double *)dev(8 * n3)) != nullptr && (d_zl = (double *)dev(8 * (size_t)(nz + 1))) != nullptr;
  if (!ok) {
    cleanup();
    return fail(c, FALL3D_ADS_EHIP, "postprocess: device allocation failed");
  }
  hipError_t e = hipSuccess;
  auto up = [&](void *d, const void *h, size_t bytes) {
    if (e == hipSuccess) e = hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, c->stream);
  };
  up(d_hm1, hm1.data(), 8 * hm1.size());
  up(d_hm2, hm2.data(), 8 * hm2.size());
  up(d_hm3, hm3.data(), 8 * hm3.size());
  up(d_spe, spe.data(), 4 * (size_t)nb);
  up(d_pmb, pmb.data(), 4 * (size_t)nb);
  if (d_zc) up(d_zc, zc.data(), 8 * n3);
  a.hm1 = d_hm1;
  a.hm2 = d_hm2;
  a.hm3 = d_hm3;
  a.spe = d_spe;
  a.pm_bin = d_pmb;
  a.zc = d_zc;
  a.con_bins = outs[0].dev;
  a.con = outs[1].dev;
  a.con_sur = outs[2].dev;
  a.col_mass = outs[3].dev;
  a.col_mass_pm = pm_ok ? outs[4].dev : nullptr;
  a.cloud_top = outs[5].dev;
  a.grn_load = outs[6].dev;
  a.grn_load_bin = outs[7].dev;
  if (e == hipSuccess) {
    (void)hipGetLastError();
    const size_t g = std::min<size_t>((ncol + 255) / 256, 65536);
    hipLaunchKernelGGL(k_output, dim3((unsigned)std::max<size_t>(g, 1)), dim3(256), 0, c->stream, a);
    e = hipGetLastError();
  }
  // cuts: the species' concentration on the z levels sigma X3max, then
  // linear interpolation to the cut positions (index and weight of each cut
  // from grid_get_shapez over the corner coordinates, on the host)
  std::vector<int> cidx[4];
  std::vector<double> csh[4], zl;
  if (cut && e == hipSuccess) {
    const fall3d_cut_args *ca = cut->args;
    zl.resize(nz + 1);
    for (int k = 0; k <= nz; ++k) zl[k] = ca->sigma[k] * ca->X3max;
    up(d_zl, zl.data(), 8 * zl.size());
    if (e == hipSuccess) {
      const size_t g = std::min<size_t>((ncol + 255) / 256, 65536);
      hipLaunchKernelGGL(k_zlevels, dim3((unsigned)std::max<size_t>(g, 1)), dim3(256), 0, c->stream, a.con, d_zc, d_zl,
                         nx, ny, nz, d_cz);
      e = hipGetLastError();
    }
    std::vector<double> xc(nx + 1), yc(ny + 1);
    for (int i = 0; i <= nx; ++i) xc[i] = ca->lonmin + (double)i * ca->dlon;
    for (int j = 0; j <= ny; ++j) yc[j] = ca->latmin + (double)j * ca->dlat;
    const int ncut[4] = {ca->ncutx, ca->ncuty, ca->ncutz, ca->nfl};
    const double *pos[4] = {ca->x_cut, ca->y_cut, ca->z_cut, ca->fl_cut};
    const std::vector<double> *coord[4] = {&xc, &yc, &zl, &zl};
    for (int q = 0; q < 4 && e == hipSuccess; ++q) {
      if (ncut[q] == 0) continue;
      cidx[q].resize(ncut[q]);
      csh[q].resize(ncut[q]);
      for (int m = 0; m < ncut[q]; ++m) {
        const double z = q == 3 ? pos[q][m] * 30.48 : pos[q][m];  // flight level (100 ft) -> m
        shapez(coord[q]->data(), 1, (int)coord[q]->size(), z, cidx[q][m], csh[q][m]);
      }
      const int mode = q < 2 ? q : 2;
      const size_t nout = (size_t)(mode == 0 ? ny + 1 : nx + 1) * (mode == 2 ? ny + 1 : nz + 1) * ncut[q];
      int *d_idx = (int *)dev(4 * (size_t)ncut[q]);
      double *d_sh = (double *)dev(8 * (size_t)ncut[q]), *d_out = (double *)dev(8 * nout);
      if (!d_idx || !d_sh || !d_out) {
        e = hipErrorOutOfMemory;
        break;
      }
      up(d_idx, cidx[q].data(), 4 * (size_t)ncut[q]);
      up(d_sh, csh[q].data(), 8 * (size_t)ncut[q]);
      if (e != hipSuccess) break;
      const size_t g = std::min<size_t>((nout + 255) / 256, 65536);
      hipLaunchKernelGGL(k_cuts, dim3((unsigned)std::max<size_t>(g, 1)), dim3(256), 0, c->stream, d_cz, nx, ny, nz,
                         mode, ncut[q], d_idx, d_sh, d_out);
      e = hipGetLastError();
      if (e == hipSuccess) e = hipMemcpyAsync(cut->out[q], d_out, 8 * nout, hipMemcpyDeviceToHost, c->stream);
    }
  }
  for (auto &x : outs)
    if (x.host && e == hipSuccess) e = hipMemcpyAsync(x.host, x.dev, 8 * x.n, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (col_mass_pm && !pm_ok && e == hipSuccess)
    for (size_t q = 0; q < outs[4].n; ++q) col_mass_pm[q] = 0.0;
  cleanup();
  if (e != hipSuccess) return fail(c, FALL3D_ADS_EHIP, "postprocess: %s", hipGetErrorString(e));
  return FALL3D_ADS_OK;
}

}  // namespace

extern "C" int fall3d_ads_postprocess(fall3d_ads_ctx *c, const fall3d_output_args *o, double *con_bins, double *con,
                                      double *con_sur, double *col_mass, double *col_mass_pm, double *cloud_top,
                                      double *grn_load, double *grn_load_bin) {
  return postprocess_impl(c, o, con_bins, con, con_sur, col_mass, col_mass_pm, cloud_top, grn_load, grn_load_bin,
                          nullptr);
}

extern "C" int fall3d_ads_postprocess_cuts(fall3d_ads_ctx *c, const fall3d_output_args *o,
                                           const fall3d_cut_args *cuts, double *cut_x, double *cut_y, double *cut_z,
                                           double *cut_fl) {
  const CutReq r = {cuts, {cut_x, cut_y, cut_z, cut_fl}};
  return postprocess_impl(c, o, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &r);
}
