// ads_enstats.hip -- ensemble statistics over member fields on gfx950
// (the statistics of Sources/task_PosEns.F90:319-441, computed per grid
// point over the members of an ensemble).
//
// Input: nmem member fields of npoints values each, member-major on the
// device (x[m * npoints + p], FP64). Outputs per point, each optional:
//   mean      sum_m x_m / nmem                       (task_PosEns.F90:321-329)
//   logmean   10^(sum_m log10 x_m / nmem)            (:334-343)
//   stddev    sqrt(sum_m (x_m - mean)^2 / (nmem-1))  (:348-365, see below)
//   median    r(int(nmem/2)), r the values in ascending order, 1-based
//                                                    (:369-383; postp_rank_vector
//                                                     sorts ascending, mod_Postp.F90:181)
//   per[i]    r(min(max(int(p_i/100 nmem), 1), nmem)) (:415-437)
//   prob[i]   100 #{m : x_m >= th_i} / nmem          (:388-410)
// The reference's standard-deviation loop assigns instead of accumulating
// (:356-360), so it would keep only the last member's squared deviation;
// this computes the sum its division by nens-1 is written for.
//
// Up to 64 members a thread owns a point and sorts its values with a
// bitonic network in registers (loads coalesced across the points); up to
// 1024 members a wave owns a point and sorts in its LDS slice.
#include <hip/hip_runtime.h>

#include <cmath>

#include "../../include/fall3d_ads.h"

namespace f3d {
namespace {

struct StatsArgs {
  const double *x;
  long npoints;
  int nmem;
  int nper, nth;
  const double *per_val;  // percentiles in percent (device)
  const double *th_val;   // thresholds (device)
  double *mean, *logmean, *stddev, *median, *per, *prob;
};

// 1-based rank of percentile p (task_PosEns.F90:429-431)
__device__ __forceinline__ int per_rank(double p, int nmem) {
  int n = (int)(p / 100.0 * nmem);
  n = n < 1 ? 1 : n;
  return n > nmem ? nmem : n;
}

// moments and exceedance of one point (member order as the reference's loops)
__device__ void point_moments(const StatsArgs &a, long p) {
  const int M = a.nmem;
  double s = 0.0, sl = 0.0;
  for (int m = 0; m < M; ++m) {
    const double v = a.x[(size_t)m * a.npoints + p];
    s = s + v;
    if (a.logmean) sl = sl + log10(v);
  }
  const double mu = s / M;
  if (a.mean) a.mean[p] = mu;
  if (a.logmean) a.logmean[p] = pow(10.0, sl / M);
  if (a.stddev) {
    double q = 0.0;
    for (int m = 0; m < M; ++m) {
      const double d = a.x[(size_t)m * a.npoints + p] - mu;
      q = q + d * d;
    }
    a.stddev[p] = M > 1 ? sqrt(q / (M - 1)) : 0.0;
  }
  for (int t = 0; t < a.nth; ++t) {
    const double th = a.th_val[t];
    int n = 0;
    for (int m = 0; m < M; ++m) n += (a.x[(size_t)m * a.npoints + p] >= th) ? 1 : 0;
    a.prob[(size_t)t * a.npoints + p] = 100.0 * n / M;
  }
}

template <int P>
__device__ __forceinline__ void bitonic_sort(double (&v)[P]) {
#pragma unroll
  for (int k = 2; k <= P; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
      for (int i = 0; i < P; ++i) {
        const int l = i ^ j;
        if (l > i) {
          const bool up = (i & k) == 0;
          const double lo = fmin(v[i], v[l]), hi = fmax(v[i], v[l]);
          v[i] = up ? lo : hi;
          v[l] = up ? hi : lo;
        }
      }
}

// selects v[r] (0-based) of a sorted register array with a runtime index
template <int P>
__device__ __forceinline__ double pick(const double (&v)[P], int r) {
  double out = v[0];
#pragma unroll
  for (int i = 1; i < P; ++i) out = (i == r) ? v[i] : out;
  return out;
}

template <int P>
__global__ void __launch_bounds__(256) k_stats_reg(StatsArgs a) {
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < a.npoints; p += (long)gridDim.x * blockDim.x) {
    point_moments(a, p);
    if (!a.median && !a.per) continue;
    double v[P];
#pragma unroll
    for (int m = 0; m < P; ++m) v[m] = (m < a.nmem) ? a.x[(size_t)m * a.npoints + p] : INFINITY;
    bitonic_sort<P>(v);
    if (a.median) a.median[p] = pick<P>(v, (a.nmem / 2 > 0 ? a.nmem / 2 : 1) - 1);
    for (int i = 0; i < a.nper; ++i) a.per[(size_t)i * a.npoints + p] = pick<P>(v, per_rank(a.per_val[i], a.nmem) - 1);
  }
}

// one wave per point, values sorted in the wave's LDS slice of P doubles
template <int P>
__global__ void __launch_bounds__(256) k_stats_wave(StatsArgs a) {
  extern __shared__ double slds[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double *v = slds + (size_t)wave * P;
  const long nw = (long)gridDim.x * (blockDim.x >> 6);
  for (long p = blockIdx.x * (long)(blockDim.x >> 6) + wave; p < a.npoints; p += nw) {
    if (lane == 0) point_moments(a, p);
    if (!a.median && !a.per) continue;
    for (int m = lane; m < P; m += 64) v[m] = (m < a.nmem) ? a.x[(size_t)m * a.npoints + p] : INFINITY;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    for (int k = 2; k <= P; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = lane; i < P; i += 64) {
          const int l = i ^ j;
          if (l > i) {
            const bool up = (i & k) == 0;
            const double x0 = v[i], x1 = v[l];
            const double lo = fmin(x0, x1), hi = fmax(x0, x1);
            v[i] = up ? lo : hi;
            v[l] = up ? hi : lo;
          }
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
      }
    if (lane == 0) {
      if (a.median) a.median[p] = v[(a.nmem / 2 > 0 ? a.nmem / 2 : 1) - 1];
      for (int i = 0; i < a.nper; ++i) a.per[(size_t)i * a.npoints + p] = v[per_rank(a.per_val[i], a.nmem) - 1];
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
}

}  // namespace
}  // namespace f3d

using namespace f3d;

extern "C" int fall3d_ens_statistics(const double *x, int nmem, long npoints, const double *percentiles, int nper,
                                     const double *thresholds, int nth, double *mean, double *logmean,
                                     double *stddev, double *median, double *per, double *prob, void *stream) {
  if (!x || nmem < 1 || nmem > 1024 || npoints < 0 || nper < 0 || nth < 0) return FALL3D_ADS_EINVAL;
  if ((nper > 0 && (!percentiles || !per)) || (nth > 0 && (!thresholds || !prob))) return FALL3D_ADS_EINVAL;
  if (npoints == 0) return FALL3D_ADS_OK;
  hipStream_t st = (hipStream_t)stream;
  StatsArgs a;
  a.x = x;
  a.npoints = npoints;
  a.nmem = nmem;
  a.nper = nper;
  a.nth = nth;
  a.per_val = percentiles;
  a.th_val = thresholds;
  a.mean = mean;
  a.logmean = logmean;
  a.stddev = stddev;
  a.median = median;
  a.per = nper > 0 ? per : nullptr;
  a.prob = nth > 0 ? prob : nullptr;
  (void)hipGetLastError();
  const long blocks_t = std::min<long>((npoints + 255) / 256, 65536);
  if (nmem <= 8)
    hipLaunchKernelGGL(k_stats_reg<8>, dim3((unsigned)blocks_t), dim3(256), 0, st, a);
  else if (nmem <= 16)
    hipLaunchKernelGGL(k_stats_reg<16>, dim3((unsigned)blocks_t), dim3(256), 0, st, a);
  else if (nmem <= 32)
    hipLaunchKernelGGL(k_stats_reg<32>, dim3((unsigned)blocks_t), dim3(256), 0, st, a);
  else if (nmem <= 64)
    hipLaunchKernelGGL(k_stats_reg<64>, dim3((unsigned)blocks_t), dim3(256), 0, st, a);
  else {
    const long blocks_w = std::min<long>((npoints + 3) / 4, 65536);
    if (nmem <= 256)
      hipLaunchKernelGGL(k_stats_wave<256>, dim3((unsigned)blocks_w), dim3(256), 4 * 256 * sizeof(double), st, a);
    else
      hipLaunchKernelGGL(k_stats_wave<1024>, dim3((unsigned)blocks_w), dim3(256), 4 * 1024 * sizeof(double), st, a);
  }
  if (hipGetLastError() != hipSuccess) return FALL3D_ADS_EHIP;
  return hipStreamSynchronize(st) == hipSuccess ? FALL3D_ADS_OK : FALL3D_ADS_EHIP;
}
