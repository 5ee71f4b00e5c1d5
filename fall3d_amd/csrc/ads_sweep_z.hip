// ads_sweep_z.hip -- instantiations of the streaming z sweep
// (Sources/mod_ADS.F90 ADS_solve_along_z); see ads_sweep_impl.h.
#include "ads_sweep_impl.h"

namespace f3d {

// Two bins per thread: Euler, and RK4 in the plain FP32 state (the
// compensated RK4 pipeline of two bins needs too many VGPRs).
#define F3D_STREAM_CASES(LAUNCH)                                                       \
  LAUNCH(1, false, 4, 1) LAUNCH(2, false, 4, 1) LAUNCH(3, false, 4, 1)                 \
  LAUNCH(1, true, 4, 1) LAUNCH(2, true, 4, 1) LAUNCH(3, true, 4, 1)                    \
  LAUNCH(1, false, 1, 1) LAUNCH(2, false, 1, 1) LAUNCH(3, false, 1, 1)                 \
  LAUNCH(1, true, 1, 1) LAUNCH(2, true, 1, 1) LAUNCH(3, true, 1, 1)                    \
  LAUNCH(1, false, 1, 2) LAUNCH(2, false, 1, 2) LAUNCH(3, false, 1, 2)                 \
  LAUNCH(1, true, 1, 2) LAUNCH(2, true, 1, 2) LAUNCH(3, true, 1, 2)                    \
  LAUNCH(1, false, 4, 2) LAUNCH(2, false, 4, 2) LAUNCH(3, false, 4, 2)

hipError_t launch_sweep_z(int bpt, int limiter, bool comp, int nst, const SweepArgs &a, dim3 grid,
                          dim3 block, hipStream_t st) {
  (void)hipGetLastError();
#define F3D_L(L, CP, N, B)                                                             \
  if (limiter == L && comp == CP && nst == N && bpt == B) {                            \
    hipLaunchKernelGGL((k_stream<2, L, CP, N, B>), grid, block, (stream_lds_bytes<2, B>(block.x)), st, a);             \
    return hipGetLastError();                                                          \
  }
  F3D_STREAM_CASES(F3D_L)
#undef F3D_L
  return hipErrorInvalidValue;
}

hipError_t launch_sweep_z_batch(int bpt, int limiter, bool comp, int nst, const SweepBatch &b, int nblocks,
                                dim3 block, hipStream_t st) {
  (void)hipGetLastError();
#define F3D_L(L, CP, N, B)                                                             \
  if (limiter == L && comp == CP && nst == N && bpt == B) {                            \
    hipLaunchKernelGGL((k_stream_batch<2, L, CP, N, B>), dim3(nblocks), block, (stream_lds_bytes<2, B>(block.x)), st, b); \
    return hipGetLastError();                                                          \
  }
  F3D_STREAM_CASES(F3D_L)
#undef F3D_L
  return hipErrorInvalidValue;
}

}  // namespace f3d
