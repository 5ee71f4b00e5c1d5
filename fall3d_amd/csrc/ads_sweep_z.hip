// ads_sweep_z.hip -- instantiations of the streaming z sweep
// (Sources/mod_ADS.F90 ADS_solve_along_z); see ads_sweep_impl.h.
#include "ads_sweep_impl.h"

namespace f3d {

hipError_t launch_sweep_z(int limiter, bool comp, int nst, const SweepArgs &a, dim3 grid, dim3 block,
                            hipStream_t st) {
  (void)hipGetLastError();
#define F3D_L(L, CP, N)                                                                     \
  if (limiter == L && comp == CP && nst == N) {                                             \
    hipLaunchKernelGGL((k_stream<2, L, CP, N>), grid, block, 0, st, a);                  \
    return hipGetLastError();                                                               \
  }
  F3D_L(1, false, 4) F3D_L(2, false, 4) F3D_L(3, false, 4)
  F3D_L(1, true, 4) F3D_L(2, true, 4) F3D_L(3, true, 4)
  F3D_L(1, false, 1) F3D_L(2, false, 1) F3D_L(3, false, 1)
  F3D_L(1, true, 1) F3D_L(2, true, 1) F3D_L(3, true, 1)
#undef F3D_L
  return hipErrorInvalidValue;
}

hipError_t launch_sweep_z_batch(int limiter, bool comp, int nst, const SweepBatch &b, int nblocks, dim3 block,
                                hipStream_t st) {
  (void)hipGetLastError();
#define F3D_L(L, CP, N)                                                                     \
  if (limiter == L && comp == CP && nst == N) {                                             \
    hipLaunchKernelGGL((k_stream_batch<2, L, CP, N>), dim3(nblocks), block, 0, st, b);     \
    return hipGetLastError();                                                               \
  }
  F3D_L(1, false, 4) F3D_L(2, false, 4) F3D_L(3, false, 4)
  F3D_L(1, true, 4) F3D_L(2, true, 4) F3D_L(3, true, 4)
  F3D_L(1, false, 1) F3D_L(2, false, 1) F3D_L(3, false, 1)
  F3D_L(1, true, 1) F3D_L(2, true, 1) F3D_L(3, true, 1)
#undef F3D_L
  return hipErrorInvalidValue;
}
}  // namespace f3d
