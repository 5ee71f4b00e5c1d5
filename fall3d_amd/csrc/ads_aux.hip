// ads_aux.hip -- the kernels around the sweeps: host-kind conversion and
// layout transposes (meteo upload), sparse source injection, deterministic
// FP64 reductions of the boundary sums, CFL time step, mass budget and the
// single-line ADS_solve_1D.
#include "ads_aux.h"

#include <climits>

namespace f3d {

// ---------------------------------------------------------------------------
// conversion / transposes
// ---------------------------------------------------------------------------

// host-kind state (dense rows of nx4 = nx+4 cells) -> (hi, lo) FP32 pair in
// rows of pitch floats (the pad words are left as they are); lo carries the
// rounding residual of an FP64 input so the compensated state starts from ~48
// significant bits. Workgroups stride over the rows, lanes along x.
template <typename T>
__global__ void k_state_in(const T *__restrict__ src, float *__restrict__ hi, float *__restrict__ lo, int nx4,
                           int pitch, size_t rows) {
  for (size_t r = blockIdx.x; r < rows; r += gridDim.x) {
    const T *s = src + r * (size_t)nx4;
    const size_t o = r * (size_t)pitch;
    for (int i = threadIdx.x; i < nx4; i += blockDim.x) {
      const double x = (double)s[i];
      const float h = (float)x;
      hi[o + i] = h;
      if (lo) lo[o + i] = (float)(x - (double)h);
    }
  }
}

template <typename T>
__global__ void k_state_out(const float *__restrict__ hi, const float *__restrict__ lo, T *__restrict__ dst, int nx4,
                            int pitch, size_t rows) {
  for (size_t r = blockIdx.x; r < rows; r += gridDim.x) {
    T *d = dst + r * (size_t)nx4;
    const size_t o = r * (size_t)pitch;
    for (int i = threadIdx.x; i < nx4; i += blockDim.x) {
      double x = (double)hi[o + i];
      if (lo) x += (double)lo[o + i];
      d[i] = (T)x;
    }
  }
}

template <typename T>
__global__ void k_to_f32(const T *__restrict__ src, float *__restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = (float)src[i];
}

template <typename T>
__global__ void k_to_f64(const T *__restrict__ src, double *__restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = (double)src[i];
}

// dst[b + B*(a + A*k)] = src[a + A*(b + B*k)]  (swap the two fastest axes),
// 32x32 tiles through LDS so both sides are coalesced.
template <typename T, typename O>
__global__ void k_transpose(const T *__restrict__ src, O *__restrict__ dst, int A, int B, int K) {
  __shared__ O tile[32][33];
  const int a0 = blockIdx.x * 32, b0 = blockIdx.y * 32;
  const int k = blockIdx.z;
  const size_t off = (size_t)A * B * k;
  for (int r = threadIdx.y; r < 32; r += blockDim.y) {
    const int a = a0 + threadIdx.x, b = b0 + r;
    if (a < A && b < B) tile[r][threadIdx.x] = (O)src[off + a + (size_t)A * b];
  }
  __syncthreads();
  for (int r = threadIdx.y; r < 32; r += blockDim.y) {
    const int b = b0 + threadIdx.x, a = a0 + r;
    if (a < A && b < B) dst[off + b + (size_t)B * a] = tile[threadIdx.x][r];
  }
}

__global__ void k_lerp(const float *__restrict__ a, const float *__restrict__ b, float *__restrict__ out, float s,
                       size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    out[i] = (1.0f - s) * a[i] + s * b[i];
}

static inline unsigned grid1d(size_t n, int bs = 256) {
  size_t g = (n + bs - 1) / bs;
  return (unsigned)(g > 65536 ? 65536 : (g == 0 ? 1 : g));
}

static inline unsigned grid_rows(size_t rows) { return (unsigned)(rows > (1u << 20) ? (1u << 20) : (rows ? rows : 1)); }

hipError_t state_in(const void *src, int kind, float *hi, float *lo, int nx4, int pitch, size_t rows, hipStream_t st) {
  (void)hipGetLastError();  // drop errors of earlier, unrelated calls
  if (kind == 8)
    hipLaunchKernelGGL(k_state_in<double>, dim3(grid_rows(rows)), dim3(256), 0, st, (const double *)src, hi, lo, nx4,
                       pitch, rows);
  else
    hipLaunchKernelGGL(k_state_in<float>, dim3(grid_rows(rows)), dim3(256), 0, st, (const float *)src, hi, lo, nx4,
                       pitch, rows);
  return hipGetLastError();
}

hipError_t state_out(const float *hi, const float *lo, void *dst, int kind, int nx4, int pitch, size_t rows,
                     hipStream_t st) {
  (void)hipGetLastError();  // drop errors of earlier, unrelated calls
  if (kind == 8)
    hipLaunchKernelGGL(k_state_out<double>, dim3(grid_rows(rows)), dim3(256), 0, st, hi, lo, (double *)dst, nx4,
                       pitch, rows);
  else
    hipLaunchKernelGGL(k_state_out<float>, dim3(grid_rows(rows)), dim3(256), 0, st, hi, lo, (float *)dst, nx4,
                       pitch, rows);
  return hipGetLastError();
}

hipError_t to_f32(const void *src, int kind, float *dst, size_t n, hipStream_t st) {
  (void)hipGetLastError();  // drop errors of earlier, unrelated calls
  if (kind == 8)
    hipLaunchKernelGGL(k_to_f32<double>, dim3(grid1d(n)), dim3(256), 0, st, (const double *)src, dst, n);
  else
    hipLaunchKernelGGL(k_to_f32<float>, dim3(grid1d(n)), dim3(256), 0, st, (const float *)src, dst, n);
  return hipGetLastError();
}

hipError_t to_f64(const void *src, int kind, double *dst, size_t n, hipStream_t st) {
  (void)hipGetLastError();  // drop errors of earlier, unrelated calls
  if (kind == 8)
    hipLaunchKernelGGL(k_to_f64<double>, dim3(grid1d(n)), dim3(256), 0, st, (const double *)src, dst, n);
  else
    hipLaunchKernelGGL(k_to_f64<float>, dim3(grid1d(n)), dim3(256), 0, st, (const float *)src, dst, n);
  return hipGetLastError();
}

// FP32 device array -> host kind (FP64 or a plain FP32 copy)
hipError_t from_f32(const float *src, void *dst, int kind, size_t n, hipStream_t st) {
  (void)hipGetLastError();  // drop errors of earlier, unrelated calls
  if (kind == 8) {
    hipLaunchKernelGGL(k_to_f64<float>, dim3(grid1d(n)), dim3(256), 0, st, src, (double *)dst, n);
    return hipGetLastError();
  }
  return hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyDeviceToDevice, st);
}

hipError_t transpose_to_f32(const void *src, int kind, float *dst, int A, int B, int K, hipStream_t st) {
  (void)hipGetLastError();  // drop errors of earlier, unrelated calls
  dim3 grid((A + 31) / 32, (B + 31) / 32, K), block(32, 8);
  if (kind == 8)
    hipLaunchKernelGGL((k_transpose<double, float>), grid, block, 0, st, (const double *)src, dst, A, B, K);
  else if (kind == 4)
    hipLaunchKernelGGL((k_transpose<float, float>), grid, block, 0, st, (const float *)src, dst, A, B, K);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t transpose_to_f64(const void *src, int kind, double *dst, int A, int B, int K, hipStream_t st) {
  (void)hipGetLastError();  // drop errors of earlier, unrelated calls
  dim3 grid((A + 31) / 32, (B + 31) / 32, K), block(32, 8);
  if (kind == 8)
    hipLaunchKernelGGL((k_transpose<double, double>), grid, block, 0, st, (const double *)src, dst, A, B, K);
  else if (kind == 4)
    hipLaunchKernelGGL((k_transpose<float, double>), grid, block, 0, st, (const float *)src, dst, A, B, K);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

template <typename T>
__global__ void k_from_f64(const double *__restrict__ src, T *__restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    dst[i] = (T)src[i];
}

hipError_t from_f64(const double *src, void *dst, int kind, size_t n, hipStream_t st) {
  (void)hipGetLastError();  // drop errors of earlier, unrelated calls
  if (kind == 8)
    hipLaunchKernelGGL(k_from_f64<double>, dim3(grid1d(n)), dim3(256), 0, st, src, (double *)dst, n);
  else
    hipLaunchKernelGGL(k_from_f64<float>, dim3(grid1d(n)), dim3(256), 0, st, src, (float *)dst, n);
  return hipGetLastError();
}

hipError_t lerp(const float *a, const float *b, float *out, float s, size_t n, hipStream_t st) {
  (void)hipGetLastError();  // drop errors of earlier, unrelated calls
  hipLaunchKernelGGL(k_lerp, dim3(grid1d(n)), dim3(256), 0, st, a, b, out, s, n);
  return hipGetLastError();
}

// dt/dxb(p) and 0.5 dt/dxb(p) in FP32 for the RK stage updates of one axis
__global__ void k_dt_over_dxb(const double *__restrict__ dxb, int n3, double dt, double wfinal,
                              float *__restrict__ tab, double *__restrict__ tab64) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n3) {
    const double f = wfinal * dt / dxb[p];
    tab[p] = (float)(dt / dxb[p]);
    tab[n3 + p] = (float)(0.5 * dt / dxb[p]);
    tab[2 * n3 + p] = (float)f;
    tab64[p] = f;
  }
}

__global__ void k_iter_table(const float *__restrict__ tab, const float *__restrict__ hdxp,
                             const double *__restrict__ dxb64, int n, int nst, float *__restrict__ itab) {
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= iter_rows(n)) return;
  const int t = row - kItabOffset;
  const int n3 = n + 3;
  auto cl = [&](int p) { return p < 0 ? 0 : (p > n + 2 ? n + 2 : p); };
  const float *dtib = tab, *hdtib = tab + n3, *ftib = tab + 2 * n3;
  float *r = itab + (size_t)row * 8;
  // stage factors are 0 outside the interior cells 1..n: a stage value at a
  // ghost position then stays the frozen ghost
  auto in = [&](int p) { return p >= 1 && p <= n; };
  r[0] = hdxp[cl(t - 2)];
  r[1] = in(t - 3) ? hdtib[t - 3] : 0.0f;
  r[2] = in(t - 5) ? hdtib[t - 5] : 0.0f;
  r[3] = in(t - 7) ? dtib[t - 7] : 0.0f;
  const int q = cl(t - 2 * nst - 1);
  r[4] = ftib[q];
  r[5] = (float)dxb64[q];  // weight of q's rounding residual
  r[6] = r[7] = 0.0f;
}

hipError_t iter_table(const float *tab, const float *hdxp, const double *dxb64, int n, int nst, float *itab,
                      hipStream_t st) {
  (void)hipGetLastError();
  const int rows = iter_rows(n);
  hipLaunchKernelGGL(k_iter_table, dim3((rows + 255) / 256), dim3(256), 0, st, tab, hdxp, dxb64, n, nst, itab);
  return hipGetLastError();
}

hipError_t dt_over_dxb(const double *dxb64, int n3, double dt, int tm, float *tab, double *tab64, hipStream_t st) {
  (void)hipGetLastError();
  const double wfinal = (tm == 1) ? 1.0 : 1.0 / 6.0;
  hipLaunchKernelGGL(k_dt_over_dxb, dim3((n3 + 255) / 256), dim3(256), 0, st, dxb64, n3, dt, wfinal, tab, tab64);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// source injection (Sources/mod_F3D.F90:1456-1473), sparse entries
// ---------------------------------------------------------------------------

__device__ __forceinline__ void comp_add2(float &hi, float &lo, float t) {
  const float s = hi + t;
  const float bp = s - hi;
  const float e = (hi - (s - bp)) + (t - bp);
  const float l = lo + e;
  const float h2 = s + l;
  lo = l - (h2 - s);
  hi = h2;
}

__device__ __forceinline__ double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// Each entry: c[idx] += dt * rate_vol (rate_vol = s/vol in FP64); the
// injected mass (increment times the cell's mass measure) is summed into
// partials[blockIdx.x*4 + 0] (slots 1..3 zero, same layout as the sweeps).
__global__ void k_source(float *__restrict__ chi, float *__restrict__ clo, const size_t *__restrict__ idx,
                         const double *__restrict__ rate_vol, const double *__restrict__ meas, int nent, double dt,
                         double *__restrict__ partials) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  double m = 0.0;
  if (e < nent) {
    const double inc = dt * rate_vol[e];
    const size_t i = idx[e];
    if (clo) {
      const float ih = (float)inc;
      const float il = (float)(inc - (double)ih);
      float h = chi[i], l = clo[i];
      comp_add2(h, l, ih);
      comp_add2(h, l, il);
      chi[i] = h;
      clo[i] = l;
      m = inc * meas[e];
    } else {
      const float old = chi[i];
      const float nv = old + (float)inc;
      chi[i] = nv;
      m = ((double)nv - (double)old) * meas[e];
    }
  }
  __shared__ double red[16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  m = wsum(m);
  if (lane == 0) red[wid] = m;
  __syncthreads();
  if (threadIdx.x < 4) {
    double s = 0.0;
    if (threadIdx.x == 0) {
      const int nw = (blockDim.x + 63) >> 6;
      for (int w = 0; w < nw; ++w) s += red[w];
    }
    partials[(size_t)blockIdx.x * 4 + threadIdx.x] = s;
  }
}

// Ensemble form: the workgroup finds its member in the prefix table of
// workgroup counts and injects that member's entries.
__global__ void k_source_batch(const SourceMember *__restrict__ tab, const int *__restrict__ bstart, int nmem) {
  int lo = 0, hi = nmem - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bstart[mid] <= (int)blockIdx.x)
      lo = mid;
    else
      hi = mid - 1;
  }
  const SourceMember sm = tab[lo];
  const int lb = blockIdx.x - bstart[lo];
  const int e = lb * blockDim.x + threadIdx.x;
  double m = 0.0;
  if (e < sm.nent) {
    const double inc = sm.dt * sm.rate_vol[e];
    const size_t i = sm.idx[e];
    if (sm.clo) {
      const float ih = (float)inc;
      const float il = (float)(inc - (double)ih);
      float h = sm.chi[i], l = sm.clo[i];
      comp_add2(h, l, ih);
      comp_add2(h, l, il);
      sm.chi[i] = h;
      sm.clo[i] = l;
      m = inc * sm.meas[e];
    } else {
      const float old = sm.chi[i];
      const float nv = old + (float)inc;
      sm.chi[i] = nv;
      m = ((double)nv - (double)old) * sm.meas[e];
    }
  }
  __shared__ double red[16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  m = wsum(m);
  if (lane == 0) red[wid] = m;
  __syncthreads();
  if (threadIdx.x < 4) {
    double s = 0.0;
    if (threadIdx.x == 0) {
      const int nw = (blockDim.x + 63) >> 6;
      for (int w = 0; w < nw; ++w) s += red[w];
    }
    sm.part[(size_t)lb * 4 + threadIdx.x] = s;
  }
}

hipError_t source_inject_batch(const SourceMember *tab, const int *bstart, int nmem, int nblocks, hipStream_t st) {
  (void)hipGetLastError();
  if (nblocks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_source_batch, dim3(nblocks), dim3(256), 0, st, tab, bstart, nmem);
  return hipGetLastError();
}

hipError_t source_inject(float *chi, float *clo, const size_t *idx, const double *rate_vol, const double *meas,
                         int nent, double dt, double *partials, int *nblocks, hipStream_t st) {
  (void)hipGetLastError();  // drop errors of earlier, unrelated calls
  const int nb = (nent + 255) / 256;
  *nblocks = nb;
  if (nb == 0) return hipSuccess;
  hipLaunchKernelGGL(k_source, dim3(nb), dim3(256), 0, st, chi, clo, idx, rate_vol, meas, nent, dt, partials);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// deterministic reduction of sweep partials into the accumulators
// ---------------------------------------------------------------------------

// One wave per (array, slot): sums partials[b*4+slot] over b in a fixed
// lane-strided order, then a fixed shuffle tree; adds into acc[dst].
// Sums of the 4 slots of p[b*4 + s] over b < nb by one 1024-thread
// workgroup, in a fixed order: thread t reads the 32-byte rows b = t + 1024 k
// (4 rows in flight, kept in separate sums and combined in order), then a
// shuffle tree per wave and the 16 wave sums in order. Valid in thread 0.
__device__ __forceinline__ void block_rows_sum(const double *__restrict__ p, int nb, double out[4]) {
  __shared__ double red[16][4];
  const int t = threadIdx.x;
  double s[4][4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int k = 0; k < 4; ++k) s[u][k] = 0.0;
  for (int b0 = 0; b0 < nb; b0 += 4096) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int b = b0 + 1024 * u + t;
      if (b < nb) {
        const double2 lo = *reinterpret_cast<const double2 *>(p + (size_t)b * 4);
        const double2 hi = *reinterpret_cast<const double2 *>(p + (size_t)b * 4 + 2);
        s[u][0] += lo.x;
        s[u][1] += lo.y;
        s[u][2] += hi.x;
        s[u][3] += hi.y;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    double v = ((s[0][k] + s[1][k]) + s[2][k]) + s[3][k];
    v = wsum(v);
    if ((t & 63) == 0) red[t >> 6][k] = v;
  }
  __syncthreads();
  if (t == 0) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      double v = 0.0;
      for (int w = 0; w < 16; ++w) v += red[w][k];
      out[k] = v;
    }
  }
}

// one workgroup (1024 threads) per job
__global__ void __launch_bounds__(1024) k_finalize(double *__restrict__ acc, ReduceJob jobs) {
  const int j = blockIdx.x;
  double s[4];
  block_rows_sum(jobs.part[j], jobs.nblocks[j], s);
  if (threadIdx.x == 0)
    for (int k = 0; k < jobs.nslot[j]; ++k) acc[jobs.dst[j][k]] += s[k];
}

// Ensemble form: one workgroup per job of a device job table.
__global__ void __launch_bounds__(1024) k_finalize_batch(const BatchJob *__restrict__ jobs) {
  const BatchJob job = jobs[blockIdx.x];
  double s[4];
  block_rows_sum(job.part, job.nblocks, s);
  if (threadIdx.x == 0)
    for (int k = 0; k < job.nslot; ++k) *job.dst[k] += s[k];
}

hipError_t finalize_batch(const BatchJob *jobs, int njobs, hipStream_t st) {
  (void)hipGetLastError();
  if (njobs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_finalize_batch, dim3(njobs), dim3(1024), 0, st, jobs);
  return hipGetLastError();
}

hipError_t finalize(double *acc, const ReduceJob &jobs, hipStream_t st) {
  (void)hipGetLastError();  // drop errors of earlier, unrelated calls
  if (jobs.count == 0) return hipSuccess;
  hipLaunchKernelGGL(k_finalize, dim3(jobs.count), dim3(1024), 0, st, acc, jobs);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// CFL time step, Sources/mod_Grid.F90:1026-1141 (FP64 arithmetic and max)
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wmax(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  return v;
}

// per-cell rates of grid_get_time_step (Sources/mod_Grid.F90:1026-1141) in
// FP64, summed in the reference's order (criterion 2: its six terms left to
// right); T is the field type of the state (float, or double for the FP64 state)
template <typename T>
__global__ void k_cfl(CflArgs a, double *__restrict__ part) {
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const T *u1 = (const T *)a.u1, *u2 = (const T *)a.u2, *v1 = (const T *)a.v1, *v2 = (const T *)a.v2;
  const T *w1 = (const T *)a.w1, *w2 = (const T *)a.w2, *k1p = (const T *)a.k1, *k2p = (const T *)a.k2;
  const T *k3p = (const T *)a.k3, *vsp = (const T *)a.vs;
  // one wave per (j, k) row at a time, lanes along x
  const int nrows = ny * nz;
  const int nw = gridDim.x * (blockDim.x >> 6);
  double m1 = 0.0, m2 = 0.0, m3 = 0.0;
  const double s = a.stime, s1 = 1.0 - a.stime;
  for (int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); r < nrows; r += nw) {
    const int k = r / ny + 1;
    const int j = r - (k - 1) * ny + 1;
    for (int i = (threadIdx.x & 63) + 1; i <= nx; i += 64) {
      const double dx = a.dx1p[i + 1], dy = a.dx2p[j + 1], dz = a.dx3p[k + 1];
      const size_t iu = (size_t)i + (size_t)(nx + 3) * ((size_t)(j - 1) + (size_t)ny * (k - 1));
      const size_t iv = (size_t)(i - 1) + (size_t)nx * ((size_t)j + (size_t)(ny + 3) * (k - 1));
      const size_t iw = (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * k);
      const double u = s1 * (double)u1[iu] + s * (double)u2[iu];
      const double v = s1 * (double)v1[iv] + s * (double)v2[iv];
      const double w = s1 * (double)w1[iw] + s * (double)w2[iw];
      const double k1 = k1p[(size_t)(i + 1) + (size_t)(nx + 4) * ((size_t)(j - 1) + (size_t)ny * (k - 1))];
      const double k2 = k2p[(size_t)(i - 1) + (size_t)nx * ((size_t)(j + 1) + (size_t)(ny + 4) * (k - 1))];
      const double k3 = k3p[(size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (k + 1))];
      double wm = 0.0;
      for (int b = 0; b < a.nb; ++b) {
        const double vs = vsp[(size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * ((size_t)k + (size_t)(nz + 3) * b))];
        wm = fmax(wm, fabs(w - vs));
      }
      const double a1 = 2.0 * k1 / (dx * dx), b1 = fabs(u) / dx;
      const double a2 = 2.0 * k2 / (dy * dy), b2 = fabs(v) / dy;
      const double a3 = 2.0 * k3 / (dz * dz), b3 = wm / dz;
      if (a.criterion == 1) {
        m1 = fmax(m1, a1 + b1);
        m2 = fmax(m2, a2 + b2);
        m3 = fmax(m3, a3 + b3);
      } else {
        m1 = fmax(m1, a1 + b1 + a2 + b2 + a3 + b3);
      }
    }
  }
  __shared__ double red[3][16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  m1 = wmax(m1);
  m2 = wmax(m2);
  m3 = wmax(m3);
  if (lane == 0) { red[0][wid] = m1; red[1][wid] = m2; red[2][wid] = m3; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r1 = 0.0, r2 = 0.0, r3 = 0.0;
    const int nw = (blockDim.x + 63) >> 6;
    for (int w = 0; w < nw; ++w) { r1 = fmax(r1, red[0][w]); r2 = fmax(r2, red[1][w]); r3 = fmax(r3, red[2][w]); }
    part[blockIdx.x] = fmax(fmax(r1, r2), r3);
  }
}

__global__ void k_cfl_final(const double *__restrict__ part, int nb, double cfl, double *__restrict__ out) {
  double m = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) m = fmax(m, part[b]);
  __shared__ double red[16];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  m = wmax(m);
  if (lane == 0) red[wid] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = 0.0;
    for (int w = 0; w < (int)((blockDim.x + 63) >> 6); ++w) r = fmax(r, red[w]);
    double dt = r;
    if (dt > 0.0) dt = 1.0 / dt;
    out[0] = cfl * dt;
  }
}

hipError_t cfl_time_step(const CflArgs &a, double *part, int part_cap, double *out, hipStream_t st) {
  (void)hipGetLastError();  // drop errors of earlier, unrelated calls
  const size_t rows = (size_t)a.ny * a.nz;  // one wave per row, 4 waves per workgroup
  if (rows > (size_t)INT_MAX) return hipErrorInvalidValue;
  int nb = (int)((rows + 3) / 4);
  if (nb > part_cap) nb = part_cap;
  if (nb < 1) nb = 1;
  if (a.fp64)
    hipLaunchKernelGGL(k_cfl<double>, dim3(nb), dim3(256), 0, st, a, part);
  else
    hipLaunchKernelGGL(k_cfl<float>, dim3(nb), dim3(256), 0, st, a, part);
  hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(256), 0, st, part, nb, a.cfl, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// mass in the volume, Sources/mod_Grid.F90:1364-1375 (FP64)
// ---------------------------------------------------------------------------

// Interior rows (j, k, bin) of the state, one wave per row at a time and the
// lanes along x: the row's position comes from one 32-bit division per row,
// and its cells are read coalesced.
struct RowWalk {
  int j, k, b;    // 1-based j, k; bin
  size_t base;    // element offset of cell (0, j, k) of bin b
};
// row r of the state with x pitch px (nx+4 for c64, state_pitch(nx) for chi/clo)
__device__ __forceinline__ RowWalk row_of(int r, int px, int ny, int nz) {
  RowWalk w;
  const int t = r / ny;
  w.j = r - t * ny + 1;
  w.k = t % nz + 1;
  w.b = t / nz;
  w.base = (size_t)1 + (size_t)px * ((size_t)(w.j + 1) + (size_t)(ny + 4) * ((size_t)(w.k + 1) + (size_t)(nz + 4) * w.b));
  return w;
}
// state value of element ci: MODE 0 FP32 (chi), 1 compensated (chi + clo), 2 FP64 (c64)
template <int MODE>
__device__ __forceinline__ double state_at(const float *chi, const float *clo, const double *c64, size_t ci) {
  if constexpr (MODE == 2) return c64[ci];
  double v = (double)chi[ci];
  if constexpr (MODE == 1) v += (double)clo[ci];
  return v;
}

// The two mass measures of the budget in one pass over the state: with the
// mass-point spacings dX_p (the reference's mass_volume; position p at index
// p+1) and with the cell sizes dX_b (index p). Per cell the products are
// c dx dy dz in that order; part[2 b + 0/1] = workgroup sums (fixed order).
template <int MODE>
__global__ void __launch_bounds__(256) k_mass(const float *__restrict__ chi, const float *__restrict__ clo,
                                              const double *__restrict__ c64, int nx, int ny, int nz, int nb,
                                              MassSpacings sp, double *__restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int nrows = ny * nz * nb;
  const int nw = gridDim.x * (blockDim.x >> 6);
  double mv = 0.0, mc = 0.0;
  for (int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); r < nrows; r += nw) {
    const RowWalk w = row_of(r, MODE == 2 ? nx + 4 : state_pitch(nx), ny, nz);
    const double pj = sp.p2[w.j + 1], pk = sp.p3[w.k + 1];
    const double bj = sp.b2[w.j], bk = sp.b3[w.k];
#pragma unroll 4
    for (int i = lane + 1; i <= nx; i += 64) {
      const double v = state_at<MODE>(chi, clo, c64, w.base + (size_t)i);
      mv += v * sp.p1[i + 1] * pj * pk;
      mc += v * sp.b1[i] * bj * bk;
    }
  }
  __shared__ double red[2][4];
  const int wid = threadIdx.x >> 6;
  mv = wsum(mv);
  mc = wsum(mc);
  if (lane == 0) {
    red[0][wid] = mv;
    red[1][wid] = mc;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double s = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) s += red[threadIdx.x][q];
    part[2 * (size_t)blockIdx.x + threadIdx.x] = s;
  }
}

// out[q] = sum over b < nblk of part[2 b + q], q = 0, 1 (fixed order)
__global__ void k_mass_final(const double *__restrict__ part, int nblk, double *__restrict__ out) {
  __shared__ double red[2][4];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) s += part[2 * (size_t)b + q];
    s = wsum(s);
    if (lane == 0) red[q][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[threadIdx.x][w];
    out[threadIdx.x] = s;
  }
}

hipError_t mass(const float *chi, const float *clo, const double *c64, int nx, int ny, int nz, int nb,
                const MassSpacings &sp, double *part, int part_cap, double *out, hipStream_t st) {
  (void)hipGetLastError();  // drop errors of earlier, unrelated calls
  const size_t rows = (size_t)ny * nz * nb;  // one wave per row, 4 waves per workgroup
  if (rows > (size_t)INT_MAX) return hipErrorInvalidValue;
  int g = (int)((rows + 3) / 4);
  if (g > part_cap / 2) g = part_cap / 2;
  if (g < 1) g = 1;
  if (c64)
    hipLaunchKernelGGL(k_mass<2>, dim3(g), dim3(256), 0, st, chi, clo, c64, nx, ny, nz, nb, sp, part);
  else if (clo)
    hipLaunchKernelGGL(k_mass<1>, dim3(g), dim3(256), 0, st, chi, clo, c64, nx, ny, nz, nb, sp, part);
  else
    hipLaunchKernelGGL(k_mass<0>, dim3(g), dim3(256), 0, st, chi, clo, c64, nx, ny, nz, nb, sp, part);
  hipLaunchKernelGGL(k_mass_final, dim3(1), dim3(256), 0, st, part, g, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// ADS_solve_1D (Sources/mod_ADS.F90:820-1053), one line in FP64, one workgroup
// ---------------------------------------------------------------------------

__device__ double lim1d(int lim, double cm, double c0, double cp) {
  const double eps = 1e-12;
  if (fabs(cp - c0) < eps) {
    if (lim == 1) return 1.0;
    if (fabs(c0 - cm) < eps) return 0.0;
    return lim == 2 ? 2.0 : 1.5;
  }
  const double r = (c0 - cm) / (cp - c0);
  if (lim == 1) return fmax(0.0, fmin(1.0, r));
  if (lim == 2) return fmax(fmax(0.0, fmin(2.0 * r, 1.0)), fmin(r, 2.0));
  return 1.5 * (r * r + r) / (r * r + r + 1.0);
}

// line arrays are indexed from the line's first halo element (position ips-2)
__global__ void k_solve_1d(Solve1DArgs a) {
  extern __shared__ double sh[];
  const int n = a.n;                 // interior cells
  double *c = sh;                    // n+4: positions -1..n+2 -> [p+1]
  double *c0 = c + (n + 4);          // n+4
  double *r = c0 + (n + 4);          // n+4
  double *kk = r + (n + 4);          // 4 stages x n
  double *flux = kk + 4 * n;         // 2
  const int t = threadIdx.x;
  for (int p = t; p < n + 4; p += blockDim.x) c[p] = a.c[p];
  __syncthreads();
  if (t == 0) {
    // boundary conditions, Sources/mod_ADS.F90:907-983
    const double ub = a.u[1], ue = a.u[n + 1];
    if (a.bcode[0] == -1) {
      if (ub > 0.0) { c[1] = c[n + 1]; c[0] = c[n]; c[n + 2] = c[2]; c[n + 3] = c[3]; }
    } else if (a.bcode[0] == 1) {
      c[1] = 2.0 * a.bvalue[0] - c[2]; c[0] = c[1];
    } else if (a.bcode[0] == 2) {
      if (ub > 0.0) { c[1] = 2.0 * a.bvalue[0] - c[2]; c[0] = c[1]; }
      else { c[1] = c[2]; c[0] = c[1]; }
    }
    if (a.bcode[1] == -1) {
      if (ue < 0.0) { c[1] = c[n + 1]; c[0] = c[n]; c[n + 2] = c[2]; c[n + 3] = c[3]; }
    } else if (a.bcode[1] == 1) {
      c[n + 2] = 2.0 * a.bvalue[1] - c[n + 1]; c[n + 3] = c[n + 2];
    } else if (a.bcode[1] == 2) {
      if (ue < 0.0) { c[n + 2] = 2.0 * a.bvalue[1] - c[n + 1]; c[n + 3] = c[n + 2]; }
      else { c[n + 2] = c[n + 1]; c[n + 3] = c[n + 2]; }
    }
  }
  __syncthreads();
  for (int p = t; p < n + 4; p += blockDim.x) c0[p] = c[p];
  __syncthreads();
  for (int st = 0; st < 4; ++st) {
    // limiter at positions 0..n+1 (array index p+1)
    for (int p = t; p <= n + 1; p += blockDim.x) r[p + 1] = lim1d(a.limiter, c0[p], c0[p + 1], c0[p + 2]);
    __syncthreads();
    if (t == 0) { r[0] = r[1]; r[n + 3] = r[n + 2]; }
    __syncthreads();
    for (int i = 1 + t; i <= n; i += blockDim.x) {
      // c_r(i), c_l(i), c_r(i+1), c_l(i+1) (Sources/mod_ADS.F90:1094-1097)
      const double cr_i = c0[i + 1] - 0.5 * r[i + 1] * (c0[i + 2] - c0[i + 1]);
      const double cl_i = c0[i] + 0.5 * r[i] * (c0[i + 1] - c0[i]);
      const double cr_p = c0[i + 2] - 0.5 * r[i + 2] * (c0[i + 3] - c0[i + 2]);
      const double cl_p = c0[i + 1] + 0.5 * r[i + 1] * (c0[i + 2] - c0[i + 1]);
      const double ui = a.u[i], up = a.u[i + 1];
      const double F_m = 0.5 * (cl_i * ui + cr_i * ui) - 0.5 * fabs(ui) * (cr_i - cl_i);
      const double F_p = 0.5 * (cl_p * up + cr_p * up) - 0.5 * fabs(up) * (cr_p - cl_p);
      const double P_m = 0.5 * (a.k[i + 1] + a.k[i]) * (c0[i + 1] - c0[i]) / a.dxp[i];
      const double P_p = 0.5 * (a.k[i + 2] + a.k[i + 1]) * (c0[i + 2] - c0[i + 1]) / a.dxp[i + 1];
      kk[st * n + i - 1] = -1.0 * (F_p - F_m) / a.dxb[i] + (P_p - P_m) / a.dxb[i];
      if (i == 1) flux[0] = F_m + P_m;
      if (i == n) flux[1] = F_p + P_p;
    }
    __syncthreads();
    if (st < 3) {
      const double h = (st == 2) ? a.dt : 0.5 * a.dt;
      for (int i = 1 + t; i <= n; i += blockDim.x) c0[i + 1] = c[i + 1] + h * kk[st * n + i - 1];
      __syncthreads();
    }
  }
  for (int i = 1 + t; i <= n; i += blockDim.x) {
    const int q = i - 1;
    c[i + 1] = c[i + 1] + a.dt * (kk[q] + 2.0 * kk[n + q] + 2.0 * kk[2 * n + q] + kk[3 * n + q]) / 6.0;
  }
  __syncthreads();
  for (int p = t; p < n + 4; p += blockDim.x) a.c[p] = c[p];
  if (t == 0) { a.flux[0] = flux[0]; a.flux[1] = flux[1]; }
}

hipError_t solve_1d(const Solve1DArgs &a, hipStream_t st) {
  (void)hipGetLastError();  // drop errors of earlier, unrelated calls
  const size_t shmem = sizeof(double) * (size_t)(3 * (a.n + 4) + 4 * a.n + 2);
  if (shmem > 60000) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_solve_1d, dim3(1), dim3(256), shmem, st, a);
  return hipGetLastError();
}

}  // namespace f3d

// ---------------------------------------------------------------------------
// Copy of known traffic for calibrating the HBM counters: every lane moves
// 16 bytes per iteration with raw buffer loads and stores, the access width
// of the sweeps' row fetches and stores. Chunks of at most 1 GiB per buffer
// descriptor (32-bit byte offsets).
// ---------------------------------------------------------------------------
namespace f3d {

__global__ void __launch_bounds__(256) k_copy_probe(const float *__restrict__ src, float *__restrict__ dst,
                                                     uint32_t n4) {
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(src), (short)0, (int)(n4 * 16u), 0x00020000);
  const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc(dst, (short)0, (int)(n4 * 16u), 0x00020000);
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n4; q += gridDim.x * blockDim.x) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)(q * 16u), 0, 0);
    __builtin_amdgcn_raw_buffer_store_b128(v, rd, (int)(q * 16u), 0, 0);
  }
}

// 8-byte accesses, the access width of the FP64 state's sweeps and transposes
// (plain global loads and stores, consecutive lanes on consecutive words)
__global__ void __launch_bounds__(256) k_copy_probe64(const double *__restrict__ src, double *__restrict__ dst,
                                                       size_t n) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    dst[q] = src[q];
}

hipError_t copy_probe64(const double *src, double *dst, size_t n, hipStream_t st) {
  (void)hipGetLastError();
  if (n) hipLaunchKernelGGL(k_copy_probe64, dim3(2048), dim3(256), 0, st, src, dst, n);
  return hipGetLastError();
}

hipError_t copy_probe(const float *src, float *dst, size_t n, hipStream_t st) {
  (void)hipGetLastError();
  constexpr size_t kChunk = (size_t)1 << 28;  // floats (1 GiB)
  for (size_t o = 0; o < n; o += kChunk) {
    const size_t m = (n - o < kChunk) ? n - o : kChunk;
    const uint32_t n4 = (uint32_t)(m / 4);
    if (n4) hipLaunchKernelGGL(k_copy_probe, dim3(2048), dim3(256), 0, st, src + o, dst + o, n4);
    for (size_t r = o + 4 * (size_t)n4; r < o + m; ++r) {  // 0-3 trailing floats
      hipError_t e = hipMemcpyAsync(dst + r, src + r, sizeof(float), hipMemcpyDeviceToDevice, st);
      if (e != hipSuccess) return e;
    }
  }
  return hipGetLastError();
}

}  // namespace f3d

// ---------------------------------------------------------------------------
// Health of the state over the interior of every bin: non-finite cells,
// negative cells and the smallest value (hi + lo for the compensated state)
// ---------------------------------------------------------------------------
namespace f3d {

template <int MODE>
__global__ void __launch_bounds__(256) k_health(const float *__restrict__ chi, const float *__restrict__ clo,
                                                const double *__restrict__ c64, int nx, int ny, int nz, int nb,
                                                double *__restrict__ part) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nrows = ny * nz * nb;
  const int nw = gridDim.x * (blockDim.x >> 6);
  double nonfin = 0.0, neg = 0.0, mn = INFINITY;
  for (int r = blockIdx.x * (blockDim.x >> 6) + wid; r < nrows; r += nw) {
    const RowWalk w = row_of(r, MODE == 2 ? nx + 4 : state_pitch(nx), ny, nz);
#pragma unroll 4
    for (int i = lane + 1; i <= nx; i += 64) {
      const double v = state_at<MODE>(chi, clo, c64, w.base + (size_t)i);
      if (!isfinite(v)) {
        nonfin += 1.0;
      } else {
        if (v < 0.0) neg += 1.0;
        mn = fmin(mn, v);
      }
    }
  }
  __shared__ double red[3][4];
  for (int o = 32; o > 0; o >>= 1) {
    nonfin += __shfl_down(nonfin, o, 64);
    neg += __shfl_down(neg, o, 64);
    mn = fmin(mn, __shfl_down(mn, o, 64));
  }
  if (lane == 0) {
    red[0][wid] = nonfin;
    red[1][wid] = neg;
    red[2][wid] = mn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, bneg = 0.0, m = INFINITY;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      a += red[0][w];
      bneg += red[1][w];
      m = fmin(m, red[2][w]);
    }
    part[blockIdx.x * 3 + 0] = a;
    part[blockIdx.x * 3 + 1] = bneg;
    part[blockIdx.x * 3 + 2] = m;
  }
}

__global__ void k_health_final(const double *__restrict__ part, int nblk, double *__restrict__ out) {
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0, m = INFINITY;
    for (int q = 0; q < nblk; ++q) {
      a += part[q * 3 + 0];
      b += part[q * 3 + 1];
      m = fmin(m, part[q * 3 + 2]);
    }
    out[0] = a;
    out[1] = b;
    out[2] = m;
  }
}

hipError_t health(const float *chi, const float *clo, const double *c64, int nx, int ny, int nz, int nb,
                  double *part, int part_cap, double *out, hipStream_t st) {
  (void)hipGetLastError();
  const size_t rows = (size_t)ny * nz * nb;  // one wave per row, 4 waves per workgroup
  if (rows > (size_t)INT_MAX) return hipErrorInvalidValue;
  int g = (int)((rows + 3) / 4);
  if (g > part_cap / 3) g = part_cap / 3;
  if (g < 1) g = 1;
  if (c64)
    hipLaunchKernelGGL(k_health<2>, dim3(g), dim3(256), 0, st, chi, clo, c64, nx, ny, nz, nb, part);
  else if (clo)
    hipLaunchKernelGGL(k_health<1>, dim3(g), dim3(256), 0, st, chi, clo, c64, nx, ny, nz, nb, part);
  else
    hipLaunchKernelGGL(k_health<0>, dim3(g), dim3(256), 0, st, chi, clo, c64, nx, ny, nz, nb, part);
  hipLaunchKernelGGL(k_health_final, dim3(1), dim3(64), 0, st, part, g, out);
  return hipGetLastError();
}

}  // namespace f3d
