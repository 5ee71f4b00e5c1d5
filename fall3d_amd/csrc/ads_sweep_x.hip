// ads_sweep_x.hip -- instantiations of the x sweep (Sources/mod_ADS.F90
// ADS_solve_along_x); see ads_sweep_impl.h.
#include "ads_sweep_impl.h"

namespace f3d {

#define F3D_CPT_LIST(X) X(2) X(3) X(4) X(6) X(8) X(12) X(16) X(20)

template <int C>
static hipError_t launch_x_c(int limiter, bool comp, int nst, const SweepArgs &a, dim3 grid, dim3 block,
                             hipStream_t st) {
  if (a.fout) {  // face-flux output: minmod, FP32 state, RK4 only (training-data runs)
    if (limiter != 1 || comp || nst != 4) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_sweep_x<C, 1, false, 4, true>), grid, block, 0, st, a);
    return hipGetLastError();
  }
#define F3D_L(L, CP, N)                                                                     \
  if (limiter == L && comp == CP && nst == N) {                                             \
    hipLaunchKernelGGL((k_sweep_x<C, L, CP, N>), grid, block, 0, st, a);                    \
    return hipGetLastError();                                                               \
  }
  F3D_L(1, false, 4) F3D_L(2, false, 4) F3D_L(3, false, 4)
  F3D_L(1, true, 4) F3D_L(2, true, 4) F3D_L(3, true, 4)
  F3D_L(1, false, 1) F3D_L(2, false, 1) F3D_L(3, false, 1)
  F3D_L(1, true, 1) F3D_L(2, true, 1) F3D_L(3, true, 1)
#undef F3D_L
  return hipErrorInvalidValue;
}

template <int C>
static hipError_t launch_xb_c(int limiter, bool comp, int nst, const SweepBatch &b, int nblocks, dim3 block,
                              hipStream_t st) {
#define F3D_L(L, CP, N)                                                                     \
  if (limiter == L && comp == CP && nst == N) {                                             \
    hipLaunchKernelGGL((k_sweep_x_batch<C, L, CP, N>), dim3(nblocks), block, 0, st, b);     \
    return hipGetLastError();                                                               \
  }
  F3D_L(1, false, 4) F3D_L(2, false, 4) F3D_L(3, false, 4)
  F3D_L(1, true, 4) F3D_L(2, true, 4) F3D_L(3, true, 4)
  F3D_L(1, false, 1) F3D_L(2, false, 1) F3D_L(3, false, 1)
  F3D_L(1, true, 1) F3D_L(2, true, 1) F3D_L(3, true, 1)
#undef F3D_L
  return hipErrorInvalidValue;
}

hipError_t launch_sweep_x_batch(int cpt, int limiter, bool comp, int nst, const SweepBatch &b, int nblocks,
                                dim3 block, hipStream_t st) {
  (void)hipGetLastError();
#define F3D_CASE(CPT) \
  if (cpt == CPT) return launch_xb_c<CPT>(limiter, comp, nst, b, nblocks, block, st);
  F3D_CPT_LIST(F3D_CASE)
#undef F3D_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_sweep_x(int cpt, int limiter, bool comp, int nst, const SweepArgs &a, dim3 grid, dim3 block,
                          hipStream_t st) {
  (void)hipGetLastError();
#define F3D_CASE(CPT) \
  if (cpt == CPT) return launch_x_c<CPT>(limiter, comp, nst, a, grid, block, st);
  F3D_CPT_LIST(F3D_CASE)
#undef F3D_CASE
  return hipErrorInvalidValue;
}

}  // namespace f3d
