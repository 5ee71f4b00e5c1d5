// ads_sweep_y.hip -- instantiations of the sweep kernel along y
// (Sources/mod_ADS.F90 ADS_solve_along_y); see ads_sweep_impl.h.
#include "ads_sweep_impl.h"

namespace f3d {

#define F3D_CPT_LIST(X) X(2) X(3) X(4) X(6) X(8) X(12) X(16)

hipError_t launch_sweep_y(int cpt, int limiter, bool comp, const SweepArgs &a, dim3 grid, dim3 block,
                            size_t shmem, hipStream_t st) {
#define F3D_CASE(CPT)                                                                              \
  if (cpt == CPT) {                                                                                \
    if (!comp) {                                                                                   \
      if (limiter == 1) { hipLaunchKernelGGL((k_sweep<1, CPT, 1, false>), grid, block, shmem, st, a); return hipGetLastError(); } \
      if (limiter == 2) { hipLaunchKernelGGL((k_sweep<1, CPT, 2, false>), grid, block, shmem, st, a); return hipGetLastError(); } \
      if (limiter == 3) { hipLaunchKernelGGL((k_sweep<1, CPT, 3, false>), grid, block, shmem, st, a); return hipGetLastError(); } \
    } else {                                                                                       \
      if (limiter == 1) { hipLaunchKernelGGL((k_sweep<1, CPT, 1, true>), grid, block, shmem, st, a); return hipGetLastError(); } \
      if (limiter == 2) { hipLaunchKernelGGL((k_sweep<1, CPT, 2, true>), grid, block, shmem, st, a); return hipGetLastError(); } \
      if (limiter == 3) { hipLaunchKernelGGL((k_sweep<1, CPT, 3, true>), grid, block, shmem, st, a); return hipGetLastError(); } \
    }                                                                                              \
    return hipErrorInvalidValue;                                                                   \
  }
  (void)hipGetLastError();
  F3D_CPT_LIST(F3D_CASE)
#undef F3D_CASE
  return hipErrorInvalidValue;
}

}  // namespace f3d
