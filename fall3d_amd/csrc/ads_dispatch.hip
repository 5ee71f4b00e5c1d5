// ads_dispatch.hip -- selects the sweep instantiation for an axis.
#include "ads_internal.h"

namespace f3d {

hipError_t launch_sweep_x(int cpt, int limiter, bool comp, int nst, const SweepArgs &a, dim3 grid, dim3 block,
                          hipStream_t st);
hipError_t launch_sweep_y(int bpt, int limiter, bool comp, int nst, const SweepArgs &a, dim3 grid, dim3 block,
                          hipStream_t st);
hipError_t launch_sweep_z(int bpt, int limiter, bool comp, int nst, const SweepArgs &a, dim3 grid, dim3 block,
                          hipStream_t st);

hipError_t launch_sweep_x_batch(int cpt, int limiter, bool comp, int nst, const SweepBatch &b, int nblocks,
                                dim3 block, hipStream_t st);
hipError_t launch_sweep_y_batch(int bpt, int limiter, bool comp, int nst, const SweepBatch &b, int nblocks,
                                dim3 block, hipStream_t st);
hipError_t launch_sweep_z_batch(int bpt, int limiter, bool comp, int nst, const SweepBatch &b, int nblocks,
                                dim3 block, hipStream_t st);

// cells per lane with an instantiated x-sweep kernel (see F3D_CPT_LIST)
const int kSweepCpt[] = {2, 3, 4, 6, 8, 12, 16, 20};
const int kSweepCptCount = 8;

bool sweep_cpt_supported(int cpt) {
  for (int i = 0; i < kSweepCptCount; ++i)
    if (kSweepCpt[i] == cpt) return true;
  return false;
}

hipError_t launch_sweep(int dir, int cpt, int limiter, bool comp, int nst, const SweepArgs &a, dim3 grid,
                        dim3 block, hipStream_t st) {
  if (dir == 0) return launch_sweep_x(cpt, limiter, comp, nst, a, grid, block, st);
  if (dir == 1) return launch_sweep_y(cpt, limiter, comp, nst, a, grid, block, st);
  if (dir == 2) return launch_sweep_z(cpt, limiter, comp, nst, a, grid, block, st);
  return hipErrorInvalidValue;
}

hipError_t launch_sweep_batch(int dir, int cpt, int limiter, bool comp, int nst, const SweepBatch &b,
                              int nblocks, dim3 block, hipStream_t st) {
  if (dir == 0) return launch_sweep_x_batch(cpt, limiter, comp, nst, b, nblocks, block, st);
  if (dir == 1) return launch_sweep_y_batch(cpt, limiter, comp, nst, b, nblocks, block, st);
  if (dir == 2) return launch_sweep_z_batch(cpt, limiter, comp, nst, b, nblocks, block, st);
  return hipErrorInvalidValue;
}

}  // namespace f3d
