// ads_mlp_device.h -- device code of the flux-correction network shared by
// the training kernels (ads_mlp.hip) and the solver's inference kernels
// (ads_learned.hip): BF16 MFMA tiles, GELU, operand layouts.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "ads_mlp.h"

namespace f3d {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

// tanh-form GELU through the identity 0.5 (1 + tanh u) = 1 / (1 + exp(-2u)):
// gelu(x) = x sigma(2u), gelu'(x) = sigma + 2 x sigma (1 - sigma) u',
// u = sqrt(2/pi) (x + 0.044715 x^3). One hardware exp and one reciprocal.
__device__ __forceinline__ float gelu_sig(float x) {
  const float u = 0.7978845608028654f * fmaf(0.044715f * x, x * x, x);
  return __fdividef(1.0f, 1.0f + __expf(-2.0f * u));
}
__device__ __forceinline__ float gelu_f(float x) { return x * gelu_sig(x); }
__device__ __forceinline__ float gelu_df(float x) {
  const float s = gelu_sig(x);
  const float du = 0.7978845608028654f * fmaf(3.0f * 0.044715f, x * x, 1.0f);
  return fmaf(2.0f * x * s * (1.0f - s), du, s);
}

// lanes of one wave exchange data through LDS: order this wave's LDS
// accesses (no workgroup barrier needed)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ f32x4 mfma(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// hidden unit of slot s, lane group g, K-chunk kc (see the header comment)
__host__ __device__ __forceinline__ int perm_h(int kc, int g, int s) {
  return 32 * kc + (s < 4 ? 4 * g + s : 16 + 4 * g + (s - 4));
}

// ---------------------------------------------------------------------------
// forward of one 16-sample column tile (all lanes of the wave)
// ---------------------------------------------------------------------------
struct TileAct {
  f32x4 z1[4], z2[4];  // pre-activations (FP32), accumulator layout
  bf16x8 hb[2];        // H1^T as B operands (bf16)
  float h2[4][4];      // GELU(z2) rounded to bf16
  float y;             // output of sample l%16 (same in the 4 lane groups)
};

__device__ __forceinline__ void load_weights(const PackedWeights *__restrict__ w, bf16x8 w1t[4], bf16x8 w2t[4][2]) {
  const int l = threadIdx.x & 63;
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
    w1t[rt] = *reinterpret_cast<const bf16x8 *>(&w->w1t[rt][l][0]);
#pragma unroll
    for (int kc = 0; kc < 2; ++kc) w2t[rt][kc] = *reinterpret_cast<const bf16x8 *>(&w->w2t[rt][kc][l][0]);
  }
}

__device__ __forceinline__ void tile_forward(const PackedWeights *__restrict__ w, const bf16x8 w1t[4],
                                             const bf16x8 w2t[4][2], bf16x8 xb, TileAct &a) {
  const int l = threadIdx.x & 63, g = l >> 4;
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
    f32x4 c;
#pragma unroll
    for (int i = 0; i < 4; ++i) c[i] = w->b1[16 * rt + 4 * g + i];
    a.z1[rt] = mfma(w1t[rt], xb, c);
  }
#pragma unroll
  for (int kc = 0; kc < 2; ++kc) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a.hb[kc][i] = (__bf16)gelu_f(a.z1[2 * kc][i]);
      a.hb[kc][4 + i] = (__bf16)gelu_f(a.z1[2 * kc + 1][i]);
    }
  }
  float y = 0.0f;
#pragma unroll
  for (int rt = 0; rt < 4; ++rt) {
    f32x4 c;
#pragma unroll
    for (int i = 0; i < 4; ++i) c[i] = w->b2[16 * rt + 4 * g + i];
    c = mfma(w2t[rt][0], a.hb[0], c);
    c = mfma(w2t[rt][1], a.hb[1], c);
    a.z2[rt] = c;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float h = (float)(__bf16)gelu_f(c[i]);
      a.h2[rt][i] = h;
      y = fmaf(w->w3[16 * rt + 4 * g + i], h, y);
    }
  }
  // the four lane groups hold disjoint hidden units of the same sample
  y += __shfl_xor(y, 16, 64);
  y += __shfl_xor(y, 32, 64);
  a.y = y + w->b3;
}

// B operand of one 16-sample tile of features: lane l reads features
// 8 (l/16) .. +7 of sample l%16 (rows beyond n read as zero)
__device__ __forceinline__ bf16x8 load_x(const uint16_t *__restrict__ x, long row, long n) {
  const int l = threadIdx.x & 63, g = l >> 4;
  bf16x8 v;
  if (row < n) {
    v = *reinterpret_cast<const bf16x8 *>(x + row * 32 + 8 * g);
  } else {
#pragma unroll
    for (int s = 0; s < 8; ++s) v[s] = (__bf16)0.0f;
  }
  return v;
}

}  // namespace f3d
