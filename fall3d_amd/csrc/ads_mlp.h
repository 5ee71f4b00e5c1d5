// ads_mlp.h -- the flux-correction network's device parameters and kernel
// launchers (ads_mlp.hip). Internal header.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace f3d {

// flat FP32 parameter vector: W1[32][64], b1[64], W2[64][64], b2[64], w3[64], b3
constexpr int MLP_NIN = 32, MLP_NH = 64;
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + 32 * 64;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + 64;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + 64 * 64;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + 64;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + 64;
constexpr int MLP_NPARAM = MLP_OFF_B3 + 1;  // 6337

// BF16 MFMA operands of the current weights (see k_pack)
struct PackedWeights {
  __bf16 w1t[4][64][8];
  __bf16 w2t[4][2][64][8];
  __bf16 w2[4][2][64][8];
  float b1[64], b2[64], w3[64];
  float b3;
  float pad[63];
};

hipError_t mlp_pack(const float *params, PackedWeights *w, hipStream_t st);
// y[r] = net(x[r]) for r < n; x: n rows of 32 bf16 (bit patterns)
hipError_t mlp_forward(const PackedWeights *w, const uint16_t *x, long n, float *y, hipStream_t st);
// waves (partial-gradient slots) used by mlp_grad for a batch of n rows
int mlp_grad_waves(long n);
// grad[0..NPARAM) = d/dparams of scale * sum_r (net(x[idx[r]]) - t[idx[r]])^2,
// grad[NPARAM] = sum of squared errors; idx may be nullptr (rows 0..n).
// part: mlp_grad_waves(n) * (NPARAM + 1) floats of scratch.
hipError_t mlp_grad(const PackedWeights *w, const uint16_t *x, const float *t, const uint32_t *idx, long n,
                    float scale, float *part, int nwaves, float *grad, hipStream_t st);
hipError_t mlp_adam(float *p, const float *grad, float *m, float *v, int step, float lr, float beta1, float beta2,
                    float eps, hipStream_t st);

}  // namespace f3d
