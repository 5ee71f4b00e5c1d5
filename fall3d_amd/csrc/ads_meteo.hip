// ads_meteo.hip -- the per-meteo-interval preprocessing of F3D_update_meteo_term
// (Sources/mod_F3D.F90:564-969) on the device: from the meteorological fields
// at the cell corners of the two bounding meteo times to the face winds, the
// diffusion coefficients and the settling velocities the sweeps read.
//
//   ground corner winds set to zero              mod_F3D.F90:662-669
//   corner -> face winds, 1 ghost face            mod_Grid.F90:685-752
//   metric scaling of the winds                   mod_Grid.F90:761-870
//   horizontal diffusion (constant, CMAQ, RAMS)   mod_Phys.F90:406-575
//   vertical diffusion (constant, similarity)     mod_Phys.F90:577-653, 1684-1741
//   settling velocity (7 drag laws, aggregates)   mod_Phys.F90:673-960
//   metric scaling of k1, k2, k3 and v_s          mod_Grid.F90:880-1017
//
// Every value is formed in FP64 with the reference's statements (the library
// is built without contraction) and stored into the sweeps' device layouts
// (ads_internal.h): FP32, and FP64 as well for the FP64 state. Each output
// element is computed by its own thread, the ghost entries from the interior
// value they copy, so no pass depends on another.
#include <cmath>

#include "ads_meteo.h"

namespace f3d {
namespace {

constexpr double kGI = 9.81;          // Sources/mod_KindType.F90:45
constexpr double kKarman = 0.4;       // :46
constexpr double kVisa0 = 1.827e-5;   // Sources/mod_Phys.F90:67
constexpr double kTa0 = 291.15;       // :68
constexpr int kCatAerosol = 2;        // Sources/mod_KindType.F90:55

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// corner arrays (1-based i in 1..nx+1, j in 1..ny+1, k in 1..nz+1)
struct Corner {
  int nx, ny;
  __device__ __forceinline__ size_t at(int i, int j, int k) const {
    return (size_t)(i - 1) + (size_t)(nx + 1) * ((size_t)(j - 1) + (size_t)(ny + 1) * (size_t)(k - 1));
  }
  __device__ __forceinline__ size_t at2(int i, int j) const { return (size_t)(i - 1) + (size_t)(nx + 1) * (j - 1); }
};

// (1 - s) a + s b, as the reference interpolates in time
__device__ __forceinline__ double lerp2(const double *a, const double *b, double s, size_t q) {
  return (1.0 - s) * a[q] + s * b[q];
}

__device__ __forceinline__ void put(float *f32, double *f64, size_t q, double v) {
  f32[q] = (float)v;
  if (f64) f64[q] = v;
}

// ---- winds ------------------------------------------------------------------
// u at x-faces f = 0..nx+2 (ghosts copy faces 1 and nx+1) divided by Hm1(j)
__global__ void k_met_u(MeteoArgs a) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t n = (size_t)(nx + 3) * ny * nz;
  const int t = blockIdx.y;
  const double *uc = a.uc[t];
  const Corner C{nx, ny};
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const int f = (int)(q % (nx + 3));
    const int j = (int)((q / (nx + 3)) % ny) + 1;
    const int k = (int)(q / ((size_t)(nx + 3) * ny)) + 1;
    const int i = clampi(f, 1, nx + 1);
    double v = 0.25 * (uc[C.at(i, j, k)] + uc[C.at(i, j + 1, k)] + uc[C.at(i, j, k + 1)] + uc[C.at(i, j + 1, k + 1)]);
    v = v / a.hm1[j - 1];
    put(a.u32[t], a.u64[t], q, v);
  }
}

// v at y-faces f = 0..ny+2 divided by Hm2 at the face, 0.5 (Hm2(f) + Hm2(f-1))
__global__ void k_met_v(MeteoArgs a) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t n = (size_t)nx * (ny + 3) * nz;
  const int t = blockIdx.y;
  const double *vc = a.vc[t];
  const Corner C{nx, ny};
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % nx) + 1;
    const int f = (int)((q / nx) % (ny + 3));
    const int k = (int)(q / ((size_t)nx * (ny + 3))) + 1;
    const int j = clampi(f, 1, ny + 1);
    double v = 0.25 * (vc[C.at(i, j, k)] + vc[C.at(i + 1, j, k)] + vc[C.at(i, j, k + 1)] + vc[C.at(i + 1, j, k + 1)]);
    const double hm2 = 0.5 * (a.hm2[f + 1] + a.hm2[f]);  // Hm2_p(p) sits at index p+1
    v = v / hm2;
    put(a.v32[t], a.v64[t], q, v);
  }
}

// w at z-faces f = 0..nz+2; sigma maps remove the terrain-following part
// (u, v at the w point times the slope) and divide by Hm3
__global__ void k_met_w(MeteoArgs a) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t n = (size_t)nx * ny * (nz + 3);
  const int t = blockIdx.y;
  const double *wc = a.wc[t], *uc = a.uc[t], *vc = a.vc[t];
  const Corner C{nx, ny};
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % nx) + 1;
    const int j = (int)((q / nx) % ny) + 1;
    const int f = (int)(q / ((size_t)nx * ny));
    const int k = clampi(f, 1, nz + 1);
    double v = 0.25 * (wc[C.at(i, j, k)] + wc[C.at(i + 1, j, k)] + wc[C.at(i, j + 1, k)] + wc[C.at(i + 1, j + 1, k)]);
    if (a.map_v != 0) {
      const size_t p2 = (size_t)(i - 1) + (size_t)nx * (j - 1);
      const double hm3 = a.hm3[p2];
      double hx = a.dhdx[p2], hy = a.dhdy[p2];
      if (a.map_v == 2) {
        const double hz = (1.0 - a.x3c[k - 1] / a.x3max);
        hx = a.dhdx[p2] * hz;
        hy = a.dhdy[p2] * hz;
      }
      const double uw = 0.25 * (uc[C.at(i, j, k)] + uc[C.at(i + 1, j, k)] + uc[C.at(i, j + 1, k)] + uc[C.at(i + 1, j + 1, k)]);
      const double vw = 0.25 * (vc[C.at(i, j, k)] + vc[C.at(i + 1, j, k)] + vc[C.at(i, j + 1, k)] + vc[C.at(i + 1, j + 1, k)]);
      v = (v - hx * uw - hy * vw) / hm3;
    }
    put(a.w32[t], a.w64[t], q, v);
  }
}

// ---- diffusion --------------------------------------------------------------
// horizontal eddy diffusivity at mass point (i, j, k), before metric scaling
__device__ double kh_point(const MeteoArgs &a, int i, int j, int k) {
  if (a.modkh == 1) return a.kh0;
  const Corner C{a.nx, a.ny};
  const double s = a.s_time;
  auto U = [&](int ii, int jj) { return lerp2(a.uc[0], a.uc[1], s, C.at(ii, jj, k)); };
  auto V = [&](int ii, int jj) { return lerp2(a.vc[0], a.vc[1], s, C.at(ii, jj, k)); };
  const double dx = a.dx1p[i + 1] * a.hm1[j - 1];  // dX1_p(i) sits at index i+1
  const double dy = a.dx2p[j + 1] * a.hm2[j + 1];
  const double dudx = 0.5 * (U(i + 1, j) - U(i, j) + U(i + 1, j + 1) - U(i, j + 1)) / dx;
  const double dvdx = 0.5 * (V(i + 1, j) - V(i, j) + V(i + 1, j + 1) - V(i, j + 1)) / dx;
  const double dudy = 0.5 * (U(i, j + 1) - U(i, j) + U(i + 1, j + 1) - U(i + 1, j)) / dy;
  const double dvdy = 0.5 * (V(i, j + 1) - V(i, j) + V(i + 1, j + 1) - V(i + 1, j)) / dy;
  constexpr double alfa = 0.28, khf = 8000.0, rkhmin = 100.0;
  if (a.modkh == 2) {  // CMAQ
    const double khn0 = a.map_h == 0 ? 4000.0 : 0.036;
    const double d1 = dudx - dvdy, d2 = dvdx + dudy;
    const double kht = alfa * alfa * dx * dy * sqrt(d1 * d1 + d2 * d2);
    const double khn = khf * (khn0 / (a.hm1[j - 1] * a.dlon)) * (khn0 / (a.hm2[j + 1] * a.dlat));
    if (kht < 1.0 && khn < 1.0) return rkhmin;
    return fmax(kht * khn / (kht + khn), rkhmin);
  }
  // RAMS (Pr = 1)
  const double kmh = 0.075 * pow(dx * dy, 2.0 / 3.0);
  const double d1 = dudy + dvdx;
  const double kk = a.rams_cs * a.rams_cs * dx * dy * sqrt(d1 * d1 + 2.0 * (dudx * dudx + dvdy * dvdy));
  return fmax(fmax(kk, kmh), rkhmin);
}

// k1 at (p = -1..nx+2, j, k): ghosts copy the first/last interior value; / Hm1(j)^2
__global__ void k_met_k1(MeteoArgs a) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t n = (size_t)(nx + 4) * ny * nz;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const int p = (int)(q % (nx + 4)) - 1;
    const int j = (int)((q / (nx + 4)) % ny) + 1;
    const int k = (int)(q / ((size_t)(nx + 4) * ny)) + 1;
    const double hm1 = a.hm1[j - 1];
    const double v = kh_point(a, clampi(p, 1, nx), j, k) / (hm1 * hm1);
    put(a.k1_32, a.k1_64, q, v);
  }
}

// k2 at (i, p = -1..ny+2, k): ghosts copy; / Hm2(p)^2
__global__ void k_met_k2(MeteoArgs a) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t n = (size_t)nx * (ny + 4) * nz;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % nx) + 1;
    const int p = (int)((q / nx) % (ny + 4)) - 1;
    const int k = (int)(q / ((size_t)nx * (ny + 4))) + 1;
    const double hm2 = a.hm2[p + 1];
    const double v = kh_point(a, i, clampi(p, 1, ny), k) / (hm2 * hm2);
    put(a.k2_32, a.k2_64, q, v);
  }
}

__device__ __forceinline__ double c2p2(const double *x, const Corner &C, int i, int j) {
  return (x[C.at2(i, j)] + x[C.at2(i + 1, j)] + x[C.at2(i, j + 1)] + x[C.at2(i + 1, j + 1)]) / 4.0;
}

// vertical eddy diffusivity at mass point (i, j, k), before metric scaling
__device__ double kv_point(const MeteoArgs &a, int i, int j, int k) {
  if (a.modkv == 1) return a.kv0;
  const Corner C{a.nx, a.ny};
  const double s = a.s_time;
  // mass-point boundary-layer fields: corner values interpolated in time, then averaged
  auto c2p2t = [&](const double *const x[2]) {
    return ((1.0 - s) * x[0][C.at2(i, j)] + s * x[1][C.at2(i, j)] + ((1.0 - s) * x[0][C.at2(i + 1, j)] + s * x[1][C.at2(i + 1, j)]) +
            ((1.0 - s) * x[0][C.at2(i, j + 1)] + s * x[1][C.at2(i, j + 1)]) +
            ((1.0 - s) * x[0][C.at2(i + 1, j + 1)] + s * x[1][C.at2(i + 1, j + 1)])) /
           4.0;
  };
  const double pblh = c2p2t(a.pblhc);
  const double ust = c2p2t(a.ustc);
  double mon = c2p2t(a.monc);
  if (fabs(mon) < 10) mon = copysign(10.0, mon);
  const double *zc = a.zc;
  const double zp = (zc[C.at(i, j, k)] + zc[C.at(i + 1, j, k)] + zc[C.at(i, j + 1, k)] + zc[C.at(i + 1, j + 1, k)] +
                     zc[C.at(i, j, k + 1)] + zc[C.at(i + 1, j, k + 1)] + zc[C.at(i, j + 1, k + 1)] +
                     zc[C.at(i + 1, j + 1, k + 1)]) /
                    8.0;
  const double hp = c2p2(a.hc, C, i, j);
  const double z = fmax(0.0, zp - hp);  // height above terrain
  double kv;
  if (z < pblh) {
    const double zl = z / mon;
    const double fih = zl < 0.0 ? 1.0 / sqrt(1.0 - 15.0 * zl) : 1.0 + 5.0 * zl;
    kv = kKarman * z * ust * (1.0 - z / pblh) * (1.0 - z / pblh) / fih;
  } else {
    constexpr double landac = 30.0;
    double lc = 1.0 / (kKarman * z) + 1.0 / landac;
    lc = 1.0 / lc;
    double z1 = 0.25 * (zc[C.at(i, j, k)] + zc[C.at(i + 1, j, k)] + zc[C.at(i, j + 1, k)] + zc[C.at(i + 1, j + 1, k)]);
    z1 = z1 - hp;
    double z2 = 0.25 * (zc[C.at(i, j, k + 1)] + zc[C.at(i + 1, j, k + 1)] + zc[C.at(i, j + 1, k + 1)] +
                        zc[C.at(i + 1, j + 1, k + 1)]);
    z2 = z2 - hp;
    auto face = [&](const double *const x[2], int kk) {
      return 0.25 * (lerp2(x[0], x[1], s, C.at(i, j, kk)) + lerp2(x[0], x[1], s, C.at(i + 1, j, kk)) +
                     lerp2(x[0], x[1], s, C.at(i, j + 1, kk)) + lerp2(x[0], x[1], s, C.at(i + 1, j + 1, kk)));
    };
    const double u1 = face(a.uc, k), v1 = face(a.vc, k);
    const double u2 = face(a.uc, k + 1), v2 = face(a.vc, k + 1);
    const double umod1 = sqrt(u1 * u1 + v1 * v1);
    const double umod2 = sqrt(u2 * u2 + v2 * v2);
    const double umod = 0.5 * (umod1 + umod2);
    const double tv1 = face(a.tvc, k), tv2 = face(a.tvc, k + 1);
    // bulk Richardson number between the two faces and its stability function
    const double um2 = fmax(umod * umod, 1e-18);
    const double ri = kGI * (z2 - z1) * (tv2 - tv1) / (um2 * 0.5 * (tv2 + tv1));
    double fc;
    if (ri > 0.0) {
      fc = 1.0 + 10.0 * ri * (1.0 + 8.0 * ri);
      fc = 1.0 / fc;
    } else {
      fc = sqrt(1.0 - 18.0 * ri);
    }
    kv = lc * lc * fabs((umod2 - umod1) / (z2 - z1)) * fc;
  }
  return fmax(0.0, kv);
}

// k3 at (i, j, p = -1..nz+2): sigma maps add the slope-projected horizontal
// diffusion and divide by Hm3^2; ghosts copy the first/last scaled value
__global__ void k_met_k3(MeteoArgs a) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t n = (size_t)nx * ny * (nz + 4);
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % nx) + 1;
    const int j = (int)((q / nx) % ny) + 1;
    const int p = (int)(q / ((size_t)nx * ny)) - 1;
    const int k = clampi(p, 1, nz);
    double v = kv_point(a, i, j, k);
    if (a.map_v != 0) {
      const size_t p2 = (size_t)(i - 1) + (size_t)nx * (j - 1);
      const double hm3 = a.hm3[p2];
      double hx = a.dhdx[p2], hy = a.dhdy[p2];
      if (a.map_v == 2) {
        const double hz = (1.0 - a.x3p[k - 1] / a.x3max);
        hx = a.dhdx[p2] * hz;
        hy = a.dhdy[p2] * hz;
      }
      const double kh = kh_point(a, i, j, k);  // k1 = k2 at the mass point before scaling
      v = (v + hx * hx * kh + hy * hy * kh) / (hm3 * hm3);
    }
    put(a.k3_32, a.k3_64, q, v);
  }
}

// ---- settling velocity ------------------------------------------------------
// terminal velocity of one particle class (Sources/mod_Phys.F90:752-960);
// returns false when the fixed-point iteration does not converge
__device__ bool vset_point(double diam, double rhop, double psi, double rhoa, double visc, int model, double &vset) {
  constexpr double atol = 1e-12, rtol = 1e-6;
  constexpr int maxit = 1000;
  double cd = 1.0;
  if (model == 4) {  // Dellino
    vset = pow((diam * diam * diam * kGI * rhop * rhoa * pow(psi, 1.6)) / (visc * visc), 0.5206);
    vset = (1.2065 * visc * vset) / (diam * rhoa);
    return true;
  }
  vset = sqrt(4.0 * kGI * diam * rhop / (3.0 * cd * rhoa));
  double vold = vset;
  // Pfeiffer's interpolation between rey = 100 and 1000
  const double cd100 = 0.24 * pow(psi, -0.828) + 2.0 * sqrt(1.07 - psi);
  const double cd1000 = 1.0;
  const double pa = (cd1000 - cd100) / 900.0;
  const double pb = cd1000 - 1000.0 * pa;
  for (int it = 1; it <= maxit; ++it) {
    const double rey = rhoa * vset * diam / visc;
    switch (model) {
      case 1:  // Arastopour
        cd = rey <= 988.947 ? 24.0 / rey * (1.0 + 0.15 * pow(rey, 0.687)) : 0.44;
        break;
      case 2: {  // Ganser
        const double rk1 = 3.0 / (1.0 + 2.0 / sqrt(psi));
        const double rk2 = pow(10.0, 1.8148 * pow(-log10(psi), 0.5743));
        cd = 24.0 / (rey * rk1) * (1.0 + 0.1118 * pow(rey * rk1 * rk2, 0.6567)) +
             0.4305 * rk2 / (1.0 + 3305.0 / (rey * rk1 * rk2));
        break;
      }
      case 3:  // Wilson
        cd = 24.0 / rey * pow(psi, -0.828) + 2.0 * sqrt(1.07 - psi);
        break;
      case 5:  // Pfeiffer
        if (rey <= 100.0)
          cd = 24.0 / rey * pow(psi, -0.828) + 2.0 * sqrt(1.07 - psi);
        else if (rey > 100.0 && rey < 1000.0)
          cd = pa * rey + pb;
        else
          cd = cd1000;
        break;
      case 6:  // Dioguardi 2017
        cd = 24.0 * (1.0 + 0.15 * pow(rey, 0.687)) / rey + 0.42 / (1.0 + 42500.0 / pow(rey, 1.16));
        cd = 0.74533 * pow(rey, 0.146012) * cd / pow(psi, 0.5134 / pow(fmax(rey, 0.01), 0.2));
        break;
      default:  // 7: Dioguardi 2018
        cd = 24.0 / rey * (pow((1.0 - psi) / rey + 1.0, 0.25) + (0.1806 * pow(rey, 0.6459)) * pow(psi, -pow(rey, 0.08))) +
             0.4251 / (1.0 + 6880.95 / rey * pow(psi, 5.03));
        break;
    }
    vset = sqrt(4.0 * kGI * diam * rhop / (3.0 * cd * rhoa));
    if (fabs(vset - vold) <= atol + rtol * fabs(vset)) return true;
    vold = vset;
  }
  return false;
}

// v_s at (i, j, f = 0..nz+2, bin): air density and temperature at the
// w-boundary from the time-interpolated corners, Sutherland viscosity,
// aggregates scaled by vset_fac, / Hm3; ghosts copy faces 1 and nz+1
// Dry deposition velocity (Feng 2008; phys_dry_deposition, Sources/mod_Phys.F90:
// 1479-1673). Coefficients a, b per land class (urban, remote continental,
// desert, polar, marine, rural) and aerosol mode (giant, coarse,
// accumulation, nuclei), :1521-1571.
__constant__ double kDryA[6][4] = {{0.0048, 0.0315, 1.2891, 1.0338}, {0.0037, 0.0120, 1.3977, 1.0707},
                                   {0.0042, 0.2928, 1.3970, 0.9155}, {0.0032, 0.1201, 1.1838, 1.0096},
                                   {0.0043, 0.1337, 1.2834, 1.1595}, {0.0045, 0.0925, 1.2654, 1.0891}};
__constant__ double kDryB[6][4] = {{1.0, 2.7925, 2.6878, 1.2644}, {1.0, 2.2413, 2.5838, 1.3247},
                                   {1.0, 3.8581, 2.5580, 1.0364}, {1.0, 3.4407, 2.8033, 1.2069},
                                   {1.0, 3.5456, 2.7157, 1.4863}, {1.0, 3.2920, 2.7227, 1.4240}};

// mode of a bin by diameter: 0 above 100 um (none), 1 giant, 2 coarse,
// 3 accumulation, 4 nuclei (:1575-1587)
__device__ __forceinline__ int drydep_mode(double d) {
  return d > 100e-6 ? 0 : d > 10e-6 ? 1 : d > 2.5e-6 ? 2 : d > 0.1e-6 ? 3 : 4;
}

// Feng land class (1..6) of a USGS 24-category land use (:1592-1608)
__device__ __forceinline__ int land_class(int cat) {
  switch (cat) {
    case 1: return 1;
    case 10: case 19: return 3;
    case 16: return 5;
    case 20: case 21: case 22: case 23: return 2;
    case 24: return 4;
    default: return 6;
  }
}

// deposition velocity of mass-point column (i, j) for a bin of the given
// mode: series of the aerodynamic resistance 1/(k u*) and the surface term
// 0.0226 u* exp(-(Re*-40300)^2 / (2 15330^2)) + a u*^b (:1613-1670)
__device__ double drydep_velocity(const MeteoArgs &a, const Corner &C, int i, int j, int mode) {
  const double s = a.s_time;
  auto U = [&](int ii, int jj) { return lerp2(a.ustc[0], a.ustc[1], s, C.at2(ii, jj)); };
  auto T = [&](int ii, int jj) { return lerp2(a.tc[0], a.tc[1], s, C.at(ii, jj, 1)); };
  auto R = [&](int ii, int jj) { return lerp2(a.rhoc[0], a.rhoc[1], s, C.at(ii, jj, 1)); };
  auto Z = [&](int ii, int jj, int k) { return a.zc[C.at(ii, jj, k)]; };
  const double ust = 0.25 * (U(i, j) + U(i + 1, j) + U(i + 1, j + 1) + U(i, j + 1));
  const double ra = ust > 0.0 ? 1.0 / (kKarman * ust) : 1e20;
  const double t = 0.25 * (T(i, j) + T(i + 1, j) + T(i + 1, j + 1) + T(i, j + 1));
  const double visc = kVisa0 * ((kTa0 + 120.0) / (t + 120.0)) * pow(t / kTa0, 1.5);
  const double rho = 0.25 * (R(i, j) + R(i + 1, j) + R(i + 1, j + 1) + R(i, j + 1));
  const double z1 = 0.25 * (Z(i, j, 1) + Z(i + 1, j, 1) + Z(i + 1, j + 1, 1) + Z(i, j + 1, 1));
  const double z2 = 0.25 * (Z(i, j, 2) + Z(i + 1, j, 2) + Z(i + 1, j + 1, 2) + Z(i, j + 1, 2));
  const double re = ust * (z2 - z1) * 0.5 * rho / visc;
  double vd1 = -0.5 * ((re - 40300.0) / 15330.0) * ((re - 40300.0) / 15330.0);
  vd1 = 0.0226 * ust * exp(vd1);
  // INT() of the land use of corner (i, j); a value outside 0..99 (NaN
  // included) is no USGS code and takes the default class, as unknown codes do
  const double lu = a.lusec[C.at2(i, j)];
  const int land = land_class(lu >= 0.0 && lu < 100.0 ? (int)lu : 0);
  double vd = vd1 + kDryA[land - 1][mode - 1] * pow(ust, kDryB[land - 1][mode - 1]);
  vd = 1.0 / (ra + 1.0 / vd);
  return vd;
}

__global__ void k_met_vs(MeteoArgs a) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const size_t n = (size_t)nx * ny * (nz + 3) * a.nb;
  const Corner C{nx, ny};
  const double s = a.s_time;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(q % nx) + 1;
    const int j = (int)((q / nx) % ny) + 1;
    const int f = (int)((q / ((size_t)nx * ny)) % (nz + 3));
    const int b = (int)(q / ((size_t)nx * ny * (nz + 3)));
    const int k = clampi(f, 1, nz + 1);
    double v = 0.0;
    if (a.cat[b] != kCatAerosol) {
      auto R = [&](int ii, int jj) { return lerp2(a.rhoc[0], a.rhoc[1], s, C.at(ii, jj, k)); };
      auto T = [&](int ii, int jj) { return lerp2(a.tc[0], a.tc[1], s, C.at(ii, jj, k)); };
      const double rhoa = 0.25 * (R(i, j) + R(i + 1, j) + R(i + 1, j + 1) + R(i, j + 1));
      const double ta = 0.25 * (T(i, j) + T(i + 1, j) + T(i + 1, j + 1) + T(i, j + 1));
      const double visa = kVisa0 * ((kTa0 + 120.0) / (ta + 120.0)) * pow(ta / kTa0, 1.5);
      if (!vset_point(a.diam[b], a.rhop[b], a.psi[b], rhoa, visa, a.modv, v)) atomicOr(a.flag, 1u);
    }
    if (a.aggregate[b]) v = a.vset_fac * v;
    if (a.drydep && f <= 1) {  // ground face and the one below it, before scaling
      const int mode = drydep_mode(a.diam[b]);
      if (mode > 0) v = v + drydep_velocity(a, C, i, j, mode);
    }
    v = v / a.hm3[(size_t)(i - 1) + (size_t)nx * (j - 1)];
    put(a.vs32, a.vs64, q, v);
  }
}

// winds of the ground corners (k = 1) of both times set to zero
__global__ void k_ground_zero(double *u0, double *u1, double *v0, double *v1, double *w0, double *w1, int nx, int ny) {
  const size_t n = (size_t)(nx + 1) * (ny + 1);
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
    u0[q] = 0.0;
    u1[q] = 0.0;
    v0[q] = 0.0;
    v1[q] = 0.0;
    w0[q] = 0.0;
    w1[q] = 0.0;
  }
}

unsigned grid_for(size_t n) {
  const size_t g = (n + 255) / 256;
  return (unsigned)(g < 1 ? 1 : (g > 65536 ? 65536 : g));
}

}  // namespace

hipError_t meteo_update(const MeteoArgs &a, hipStream_t st) {
  (void)hipGetLastError();
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  hipLaunchKernelGGL(k_ground_zero, dim3(grid_for((size_t)(nx + 1) * (ny + 1))), dim3(256), 0, st, a.uc[0], a.uc[1],
                     a.vc[0], a.vc[1], a.wc[0], a.wc[1], nx, ny);
  hipLaunchKernelGGL(k_met_u, dim3(grid_for((size_t)(nx + 3) * ny * nz), 2), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_met_v, dim3(grid_for((size_t)nx * (ny + 3) * nz), 2), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_met_w, dim3(grid_for((size_t)nx * ny * (nz + 3)), 2), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_met_k1, dim3(grid_for((size_t)(nx + 4) * ny * nz)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_met_k2, dim3(grid_for((size_t)nx * (ny + 4) * nz)), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_met_k3, dim3(grid_for((size_t)nx * ny * (nz + 4))), dim3(256), 0, st, a);
  hipLaunchKernelGGL(k_met_vs, dim3(grid_for((size_t)nx * ny * (nz + 3) * a.nb)), dim3(256), 0, st, a);
  return hipGetLastError();
}

}  // namespace f3d

// =============================================================================
// C ABI
// =============================================================================
#include <cstring>
#include <vector>

#include "ads_aux.h"
#include "ads_context.h"

namespace f3d {
namespace host {
namespace {

// device FP64 arrays of one update, freed together
struct DevArrays {
  std::vector<void *> owned;
  ~DevArrays() {
    for (void *p : owned) dfree(p);
  }
};

// host array of the context kind (n values) -> new FP64 device array; raw
// is a device buffer of at least n * kind bytes for the transfer
int up64(fall3d_ads_ctx *c, DevArrays &d, const void *h, size_t n, void *raw, double **out) {
  *out = nullptr;
  double *p = nullptr;
  HIPCHK(c, dalloc(&p, n));
  d.owned.push_back(p);
  HIPCHK(c, hipMemcpyAsync(raw, h, n * (size_t)c->kind, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, to_f64(raw, c->kind, p, n, c->stream));
  // the transfer buffer is reused by the next upload: finish this one first
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *out = p;
  return FALL3D_ADS_OK;
}

template <typename T>
int up_bins(fall3d_ads_ctx *c, DevArrays &d, const T *h, int n, T **out) {
  T *p = nullptr;
  HIPCHK(c, dalloc(&p, (size_t)n));
  d.owned.push_back(p);
  HIPCHK(c, hipMemcpy(p, h, sizeof(T) * n, hipMemcpyHostToDevice));
  *out = p;
  return FALL3D_ADS_OK;
}

// device array (FP32 or FP64) of n values -> host doubles
int down64(fall3d_ads_ctx *c, const float *d32, const double *d64, size_t n, std::vector<double> &out) {
  out.resize(n);
  if (d64) {
    HIPCHK(c, hipMemcpy(out.data(), d64, n * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::vector<float> t(n);
    HIPCHK(c, hipMemcpy(t.data(), d32, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t q = 0; q < n; ++q) out[q] = (double)t[q];
  }
  return FALL3D_ADS_OK;
}

void put_host(void *h, int kind, size_t q, double v) {
  if (kind == 8)
    ((double *)h)[q] = v;
  else
    ((float *)h)[q] = (float)v;
}

// device layout dev[a + A*(b + B*r)] -> host layout h[b + B*(a + A*r)] (the
// two fastest axes swapped back), A*B*R values
void unswap(const std::vector<double> &dev, void *h, int kind, size_t A, size_t B, size_t R) {
  for (size_t r = 0; r < R; ++r)
    for (size_t a = 0; a < A; ++a)
      for (size_t b = 0; b < B; ++b) put_host(h, kind, b + B * (a + A * r), dev[a + A * (b + B * r)]);
}

}  // namespace
}  // namespace host
}  // namespace f3d

using namespace f3d;
using namespace f3d::host;

extern "C" {

int fall3d_ads_update_meteo(fall3d_ads_ctx *c, const fall3d_meteo_corner *m, const fall3d_meteo_grid *g,
                            const fall3d_meteo_options *o) {
  if (!c) return FALL3D_ADS_EINVAL;
  if (!m || !g || !o) return fail(c, FALL3D_ADS_EINVAL, "null argument");
  if (!c->grid_ok) return fail(c, FALL3D_ADS_ESTATE, "fall3d_ads_set_grid has not been called");
  if (o->map_h != 0 && o->map_h != 1) return fail(c, FALL3D_ADS_EINVAL, "horizontal mapping %d not implemented", o->map_h);
  if (o->map_v < 0 || o->map_v > 2) return fail(c, FALL3D_ADS_EINVAL, "vertical mapping %d not implemented", o->map_v);
  if (o->modkh != 1 && o->modkh != 2 && o->modkh != 4)
    return fail(c, FALL3D_ADS_EINVAL, "horizontal diffusion model %d not implemented", o->modkh);
  if (o->modkv != 1 && o->modkv != 3)
    return fail(c, FALL3D_ADS_EINVAL, "vertical diffusion model %d not implemented", o->modkv);
  if (o->modv < 1 || o->modv > 7) return fail(c, FALL3D_ADS_EINVAL, "settling velocity model %d not found", o->modv);
  if (o->nbins != c->nb) return fail(c, FALL3D_ADS_EINVAL, "%d bins given, context has %d", o->nbins, c->nb);
  if (!o->diam || !o->rho || !o->psi || !o->cat || !o->aggregate)
    return fail(c, FALL3D_ADS_EINVAL, "null bin property array");
  const bool sim = o->modkv == 3;
  for (int t = 0; t < 2; ++t) {
    if (!m->uc[t] || !m->vc[t] || !m->wc[t] || !m->tc[t] || !m->rhoc[t])
      return fail(c, FALL3D_ADS_EINVAL, "null corner field");
    if (sim && (!m->tvc[t] || !m->pblhc[t] || !m->ustc[t] || !m->monc[t]))
      return fail(c, FALL3D_ADS_EINVAL, "similarity needs tvc, pblhc, ustc and monc");
  }
  if (!g->Hm1_p || !g->Hm2_p || !g->Hm3_p) return fail(c, FALL3D_ADS_EINVAL, "null map factor array");
  if (o->map_v != 0 && (!g->dhdx_p || !g->dhdy_p)) return fail(c, FALL3D_ADS_EINVAL, "sigma mapping needs dhdx_p, dhdy_p");
  if (o->map_v == 2 && (!g->X3_c || !g->X3_p || !(o->x3max > 0.0)))
    return fail(c, FALL3D_ADS_EINVAL, "linear decay needs X3_c, X3_p and x3max > 0");
  if (sim && (!g->z_c || !g->h_c)) return fail(c, FALL3D_ADS_EINVAL, "similarity needs z_c and h_c");
  const bool dd = o->dry_deposition != 0;
  if (dd && (!m->ustc[0] || !m->ustc[1] || !m->lusec || !g->z_c))
    return fail(c, FALL3D_ADS_EINVAL, "dry deposition needs ustc, lusec and z_c");

  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const size_t n3 = (size_t)(nx + 1) * (ny + 1) * (nz + 1), n2 = (size_t)(nx + 1) * (ny + 1);
  DevArrays d;
  void *raw = nullptr;
  HIPCHK(c, dalloc((char **)&raw, n3 * (size_t)c->kind));
  d.owned.push_back(raw);
  MeteoArgs a;
  memset(&a, 0, sizeof(a));
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  a.nb = c->nb;
  a.map_h = o->map_h;
  a.map_v = o->map_v;
  a.modkh = o->modkh;
  a.modkv = o->modkv;
  a.modv = o->modv;
  a.kh0 = o->kh0;
  a.kv0 = o->kv0;
  a.rams_cs = o->rams_cs;
  a.dlon = o->dlon;
  a.dlat = o->dlat;
  a.x3max = o->x3max;
  a.s_time = o->s_time;
  a.vset_fac = o->vset_fac;
  int rc;
  double *tmp = nullptr;
  for (int t = 0; t < 2; ++t) {
    if ((rc = up64(c, d, m->uc[t], n3, raw, &a.uc[t]))) return rc;
    if ((rc = up64(c, d, m->vc[t], n3, raw, &a.vc[t]))) return rc;
    if ((rc = up64(c, d, m->wc[t], n3, raw, &a.wc[t]))) return rc;
    if ((rc = up64(c, d, m->tc[t], n3, raw, &tmp))) return rc;
    a.tc[t] = tmp;
    if ((rc = up64(c, d, m->rhoc[t], n3, raw, &tmp))) return rc;
    a.rhoc[t] = tmp;
    if (sim) {
      if ((rc = up64(c, d, m->tvc[t], n3, raw, &tmp))) return rc;
      a.tvc[t] = tmp;
      if ((rc = up64(c, d, m->pblhc[t], n2, raw, &tmp))) return rc;
      a.pblhc[t] = tmp;
      if ((rc = up64(c, d, m->monc[t], n2, raw, &tmp))) return rc;
      a.monc[t] = tmp;
    }
    if (sim || dd) {
      if ((rc = up64(c, d, m->ustc[t], n2, raw, &tmp))) return rc;
      a.ustc[t] = tmp;
    }
  }
  if ((rc = up64(c, d, g->Hm1_p, (size_t)ny, raw, &tmp))) return rc;
  a.hm1 = tmp;
  if ((rc = up64(c, d, g->Hm2_p, (size_t)ny + 4, raw, &tmp))) return rc;
  a.hm2 = tmp;
  if ((rc = up64(c, d, g->Hm3_p, (size_t)nx * ny, raw, &tmp))) return rc;
  a.hm3 = tmp;
  if (o->map_v != 0) {
    if ((rc = up64(c, d, g->dhdx_p, (size_t)nx * ny, raw, &tmp))) return rc;
    a.dhdx = tmp;
    if ((rc = up64(c, d, g->dhdy_p, (size_t)nx * ny, raw, &tmp))) return rc;
    a.dhdy = tmp;
  }
  if (o->map_v == 2) {
    if ((rc = up64(c, d, g->X3_c, (size_t)nz + 1, raw, &tmp))) return rc;
    a.x3c = tmp;
    if ((rc = up64(c, d, g->X3_p, (size_t)nz, raw, &tmp))) return rc;
    a.x3p = tmp;
  }
  if (sim || dd) {
    if ((rc = up64(c, d, g->z_c, n3, raw, &tmp))) return rc;
    a.zc = tmp;
  }
  if (sim) {
    if ((rc = up64(c, d, g->h_c, n2, raw, &tmp))) return rc;
    a.hc = tmp;
  }
  if (dd) {
    if ((rc = up64(c, d, m->lusec, n2, raw, &tmp))) return rc;
    a.lusec = tmp;
    a.drydep = 1;
  }
  a.dx1p = c->dxp64[0];
  a.dx2p = c->dxp64[1];
  double *bd = nullptr, *br = nullptr, *bp = nullptr;
  int *bc = nullptr, *ba = nullptr;
  if ((rc = up_bins(c, d, o->diam, c->nb, &bd))) return rc;
  if ((rc = up_bins(c, d, o->rho, c->nb, &br))) return rc;
  if ((rc = up_bins(c, d, o->psi, c->nb, &bp))) return rc;
  if ((rc = up_bins(c, d, o->cat, c->nb, &bc))) return rc;
  if ((rc = up_bins(c, d, o->aggregate, c->nb, &ba))) return rc;
  a.diam = bd;
  a.rhop = br;
  a.psi = bp;
  a.cat = bc;
  a.aggregate = ba;
  unsigned *flag = nullptr;
  HIPCHK(c, dalloc(&flag, 1));
  d.owned.push_back(flag);
  HIPCHK(c, hipMemsetAsync(flag, 0, sizeof(unsigned), c->stream));
  a.flag = flag;
  a.u32[0] = c->u1;
  a.u32[1] = c->u2;
  a.v32[0] = c->v1;
  a.v32[1] = c->v2;
  a.w32[0] = c->w1;
  a.w32[1] = c->w2;
  a.k1_32 = c->k1;
  a.k2_32 = c->k2;
  a.k3_32 = c->k3;
  a.vs32 = c->vs;
  a.u64[0] = c->u1d;
  a.u64[1] = c->u2d;
  a.v64[0] = c->v1d;
  a.v64[1] = c->v2d;
  a.w64[0] = c->w1d;
  a.w64[1] = c->w2d;
  a.k1_64 = c->k1d;
  a.k2_64 = c->k2d;
  a.k3_64 = c->k3d;
  a.vs64 = c->vsd;
  {
    hipError_t e = meteo_update(a, c->stream);
    if (e != hipSuccess) return fail(c, FALL3D_ADS_EHIP, "meteo update: %s", hipGetErrorString(e));
  }
  unsigned hflag = 0;
  HIPCHK(c, hipMemcpyAsync(&hflag, flag, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (hflag & 1u) return fail(c, FALL3D_ADS_ESTATE, "no convergence in vset");
  c->meteo_ok = true;
  return FALL3D_ADS_OK;
}

int fall3d_ads_get_meteo(fall3d_ads_ctx *c, void *u1, void *u2, void *v1, void *v2, void *w1, void *w2, void *k1,
                         void *k2, void *k3, void *vs) {
  if (!c) return FALL3D_ADS_EINVAL;
  if (!c->meteo_ok) return fail(c, FALL3D_ADS_ESTATE, "the context has no meteo");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t nx = c->nx, ny = c->ny, nz = c->nz, nb = c->nb;
  const bool f64 = is_fp64(c);
  std::vector<double> t;
  int rc;
  auto same = [&](void *h, const float *d32, const double *d64, size_t n) -> int {
    if (!h) return FALL3D_ADS_OK;
    int r = down64(c, d32, f64 ? d64 : nullptr, n, t);
    if (r) return r;
    for (size_t q = 0; q < n; ++q) put_host(h, c->kind, q, t[q]);
    return FALL3D_ADS_OK;
  };
  auto swapped = [&](void *h, const float *d32, const double *d64, size_t A, size_t B, size_t R) -> int {
    if (!h) return FALL3D_ADS_OK;
    int r = down64(c, d32, f64 ? d64 : nullptr, A * B * R, t);
    if (r) return r;
    unswap(t, h, c->kind, A, B, R);
    return FALL3D_ADS_OK;
  };
  // device: u (nx+3, ny, nz), v (nx, ny+3, nz), w (nx*ny, nz+3), k1 (nx+4, ny, nz),
  // k2 (nx, ny+4, nz), k3 (nx*ny, nz+4), vs (nx, ny, nz+3, nb): x fastest
  if ((rc = same(u1, c->u1, c->u1d, (nx + 3) * ny * nz))) return rc;
  if ((rc = same(u2, c->u2, c->u2d, (nx + 3) * ny * nz))) return rc;
  if ((rc = swapped(v1, c->v1, c->v1d, nx, ny + 3, nz))) return rc;
  if ((rc = swapped(v2, c->v2, c->v2d, nx, ny + 3, nz))) return rc;
  if ((rc = swapped(w1, c->w1, c->w1d, nx * ny, nz + 3, 1))) return rc;
  if ((rc = swapped(w2, c->w2, c->w2d, nx * ny, nz + 3, 1))) return rc;
  if ((rc = same(k1, c->k1, c->k1d, (nx + 4) * ny * nz))) return rc;
  if ((rc = swapped(k2, c->k2, c->k2d, nx, ny + 4, nz))) return rc;
  if ((rc = swapped(k3, c->k3, c->k3d, nx * ny, nz + 4, 1))) return rc;
  if ((rc = same(vs, c->vs, c->vsd, nx * ny * (nz + 3) * nb))) return rc;
  return FALL3D_ADS_OK;
}

}  // extern "C"
