// ads_mlp.hip -- the per-face flux-correction network on gfx950 MFMA.
//
// Network (no counterpart in the FALL3D reference; the learned part of the
// coarse-grid solver): 32 features -> 64 -> 64 -> 1, tanh-form GELU, MSE
// loss, Adam. Weights and activations are BF16, every product accumulates
// in FP32 (v_mfma_f32_16x16x32_bf16), the master weights and the Adam
// moments are FP32.
//
// Layout trick: layers are computed transposed, Z^T = W^T X^T, with
// samples along the MFMA's N dimension. The accumulator layout of one
// 16x16 tile (lane l holds rows 4(l/16)+i, i=0..3, of column l%16) then
// feeds the next layer's B operand (lane l holds 8 consecutive k of column
// l%16) without any data movement, provided the k index of that product is
// permuted: slot s of lane group g in K-chunk kc is hidden unit
// 32 kc + (s < 4 ? 4 g + s : 16 + 4 g + s - 4). The A operands (weights)
// are packed in the same permuted order once per update (k_pack), so the
// permutation costs nothing at run time. Weight gradients contract over
// samples and need the transposed arrangement: those tiles go through LDS.
//
// Parameter vector (FP32, shared with the FP64 oracle):
//   W1[32][64] (k-major), b1[64], W2[64][64] (in-major), b2[64], w3[64], b3.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "ads_mlp.h"
#include "ads_mlp_device.h"

namespace f3d {

// ---------------------------------------------------------------------------
// packing of the FP32 master weights into the MFMA operand layouts
// ---------------------------------------------------------------------------
// PackedWeights (device): per lane (64) and tile, 8 bf16:
//   w1t[rt][lane]  : A of Z1^T = W1^T X^T : row h = 16 rt + l%16, k = 8 (l/16) + s -> W1[k][h]
//   w2t[rt][kc][l] : A of Z2^T = W2^T H1^T: row o = 16 rt + l%16, k = perm_h(kc, l/16, s) -> W2[k][o]
//   w2[rt][kc][l]  : A of dH1^T = W2 dZ2^T: row i = 16 rt + l%16, k = perm_h(kc, l/16, s) -> W2[i][k]
//   b1, b2, w3 (rounded to bf16, stored as float), b3
__global__ void k_pack(const float *__restrict__ p, PackedWeights *__restrict__ w) {
  const int l = threadIdx.x & 63, g = l >> 4, c = l & 15;
  const float *W1 = p + MLP_OFF_W1, *W2 = p + MLP_OFF_W2;
  for (int rt = 0; rt < 4; ++rt) {
    for (int s = 0; s < 8; ++s) {
      w->w1t[rt][l][s] = (__bf16)W1[(8 * g + s) * 64 + 16 * rt + c];
      for (int kc = 0; kc < 2; ++kc) {
        const int h = perm_h(kc, g, s);
        w->w2t[rt][kc][l][s] = (__bf16)W2[h * 64 + 16 * rt + c];
        w->w2[rt][kc][l][s] = (__bf16)W2[(16 * rt + c) * 64 + h];
      }
    }
  }
  w->b1[l] = (float)(__bf16)p[MLP_OFF_B1 + l];
  w->b2[l] = (float)(__bf16)p[MLP_OFF_B2 + l];
  w->w3[l] = (float)(__bf16)p[MLP_OFF_W3 + l];
  if (l == 0) w->b3 = (float)(__bf16)p[MLP_OFF_B3];
}

__global__ void __launch_bounds__(256) k_mlp_forward(const PackedWeights *__restrict__ w,
                                                     const uint16_t *__restrict__ x, long n,
                                                     float *__restrict__ y) {
  const int l = threadIdx.x & 63;
  bf16x8 w1t[4], w2t[4][2];
  load_weights(w, w1t, w2t);
  const long ntile = (n + 15) / 16;
  const long wave = (long)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const long nwave = (long)gridDim.x * (blockDim.x >> 6);
  for (long t = wave; t < ntile; t += nwave) {
    const long row = t * 16 + (l & 15);
    TileAct a;
    tile_forward(w, w1t, w2t, load_x(x, row, n), a);
    if (l < 16 && row < n) y[row] = a.y;
  }
}

// ---------------------------------------------------------------------------
// forward + backward over a batch: per-wave partial gradients (FP32)
// ---------------------------------------------------------------------------
// Wave w takes the 32 rows r = 32 w .. 32 w + 31 of the batch (row idx[r],
// or r when idx == nullptr; rows >= n contribute nothing) as two 16-sample
// column tiles, and writes NPARAM + 1 floats to part[w]: the gradient of
// scale * sum (y - t)^2 over its rows, then their sum of squared errors.
// The weight gradients contract over samples: H1^T, dZ2^T, dZ1^T and X^T
// are staged in LDS (bf16) and multiplied with MFMA; the bias gradients
// are row sums of the same staged tiles.
__global__ void __launch_bounds__(256) k_mlp_grad(const PackedWeights *__restrict__ w,
                                                  const uint16_t *__restrict__ x, const float *__restrict__ tgt,
                                                  const uint32_t *__restrict__ idx, long n, float scale,
                                                  float *__restrict__ part) {
  // per wave: H1^T [64][32], dZ2^T [64][32], dZ1^T [64][32], X^T [32][32] (bf16)
  __shared__ __bf16 lds[4][(64 + 64 + 64 + 32) * 32];
  const int l = threadIdx.x & 63, g = l >> 4, c = l & 15, wv = threadIdx.x >> 6;
  __bf16 *sh1 = lds[wv], *sdz2 = sh1 + 64 * 32, *sdz1 = sdz2 + 64 * 32, *sx = sdz1 + 64 * 32;
  const long wave = (long)blockIdx.x * (blockDim.x >> 6) + wv;

  float gw3[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int i = 0; i < 4; ++i) gw3[a][i] = 0.0f;
  float gb3 = 0.0f, sse = 0.0f;
  {
    bf16x8 w1t[4], w2t[4][2];
    load_weights(w, w1t, w2t);
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
      const long r = wave * 32 + ct * 16 + c;  // position in the batch
      const bool ok = r < n;
      const long row = ok ? (idx ? (long)idx[r] : r) : 0;
      const bf16x8 xb = load_x(x, row, ok ? row + 1 : 0);
      TileAct a;
      tile_forward(w, w1t, w2t, xb, a);
      const float res = ok ? a.y - tgt[row] : 0.0f;
      const float dy = 2.0f * scale * res;
      if (g == 0) {
        sse = fmaf(res, res, sse);
        gb3 += dy;
      }
      // dZ2 = w3 dy gelu'(z2) (bf16, B-operand order) and dw3 = sum h2 dy
      bf16x8 dzb[2];
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          gw3[rt][i] = fmaf(a.h2[rt][i], dy, gw3[rt][i]);
          const __bf16 d = (__bf16)(w->w3[16 * rt + 4 * g + i] * dy * gelu_df(a.z2[rt][i]));
          if (rt & 1)
            dzb[rt >> 1][4 + i] = d;
          else
            dzb[rt >> 1][i] = d;
        }
      // dH1^T = W2 dZ2^T ; dZ1 = dH1 gelu'(z1), staged as dZ1^T
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) {
        const bf16x8 A0 = *reinterpret_cast<const bf16x8 *>(&w->w2[rt][0][l][0]);
        const bf16x8 A1 = *reinterpret_cast<const bf16x8 *>(&w->w2[rt][1][l][0]);
        f32x4 d = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        d = mfma(A0, dzb[0], d);
        d = mfma(A1, dzb[1], d);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int h = 16 * rt + 4 * g + i;
          sdz1[h * 32 + ct * 16 + c] = (__bf16)(d[i] * gelu_df(a.z1[rt][i]));
        }
      }
      // H1^T and dZ2^T (rows = hidden unit, columns = sample), X^T
#pragma unroll
      for (int kc = 0; kc < 2; ++kc)
#pragma unroll
        for (int s2 = 0; s2 < 8; ++s2) {
          const int h = perm_h(kc, g, s2);
          sh1[h * 32 + ct * 16 + c] = a.hb[kc][s2];
          sdz2[h * 32 + ct * 16 + c] = dzb[kc][s2];
        }
#pragma unroll
      for (int s2 = 0; s2 < 8; ++s2) sx[(8 * g + s2) * 32 + ct * 16 + c] = xb[s2];
    }
  }
  wave_sync();  // the tiles were written by other lanes of this wave

  float *out = part + (size_t)wave * (MLP_NPARAM + 1);
  // dW2^T[o][i] = sum_s dZ2^T[o][s] H1^T[i][s]   (A: row o, k = s; B: k = s, col i)
#pragma unroll
  for (int ot = 0; ot < 4; ++ot) {
    const bf16x8 A = *reinterpret_cast<const bf16x8 *>(sdz2 + (16 * ot + c) * 32 + 8 * g);
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const bf16x8 B = *reinterpret_cast<const bf16x8 *>(sh1 + (16 * it + c) * 32 + 8 * g);
      const f32x4 acc = mfma(A, B, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
      for (int i = 0; i < 4; ++i) out[MLP_OFF_W2 + (16 * it + c) * 64 + 16 * ot + 4 * g + i] = acc[i];
    }
  }
  // dW1^T[h][k] = sum_s dZ1^T[h][s] X^T[k][s]
#pragma unroll
  for (int ht = 0; ht < 4; ++ht) {
    const bf16x8 A = *reinterpret_cast<const bf16x8 *>(sdz1 + (16 * ht + c) * 32 + 8 * g);
#pragma unroll
    for (int kt = 0; kt < 2; ++kt) {
      const bf16x8 B = *reinterpret_cast<const bf16x8 *>(sx + (16 * kt + c) * 32 + 8 * g);
      const f32x4 acc = mfma(A, B, f32x4{0.0f, 0.0f, 0.0f, 0.0f});
#pragma unroll
      for (int i = 0; i < 4; ++i) out[MLP_OFF_W1 + (16 * kt + c) * 64 + 16 * ht + 4 * g + i] = acc[i];
    }
  }
  // db1[h], db2[h]: row sums of the staged dZ tiles (lane h)
  {
    float s1 = 0.0f, s2 = 0.0f;
#pragma unroll
    for (int q = 0; q < 32; q += 8) {
      const bf16x8 v1 = *reinterpret_cast<const bf16x8 *>(sdz1 + l * 32 + q);
      const bf16x8 v2 = *reinterpret_cast<const bf16x8 *>(sdz2 + l * 32 + q);
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        s1 += (float)v1[e];
        s2 += (float)v2[e];
      }
    }
    out[MLP_OFF_B1 + l] = s1;
    out[MLP_OFF_B2 + l] = s2;
  }
  // dw3: reduce over the 16 lanes (samples) of each group
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float v = gw3[rt][i];
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
      if (c == 0) out[MLP_OFF_W3 + 16 * rt + 4 * g + i] = v;
    }
  for (int o = 1; o < 64; o <<= 1) {
    gb3 += __shfl_xor(gb3, o, 64);
    sse += __shfl_xor(sse, o, 64);
  }
  if (l == 0) {
    out[MLP_OFF_B3] = gb3;
    out[MLP_NPARAM] = sse;
  }
}

// grad[p] = sum over waves of part[w][p], p <= NPARAM, in a fixed order: a
// workgroup takes 64 parameters; its 16 waves sum the partials w = r,
// r + 16, r + 32, ... (r = wave), then wave 0 adds the 16 sums in order r.
constexpr int kReduceSlots = 16;
__global__ void __launch_bounds__(64 * kReduceSlots) k_mlp_reduce(const float *__restrict__ part, int nwave,
                                                                  float *__restrict__ grad) {
  __shared__ float sums[kReduceSlots][64];
  const int l = threadIdx.x & 63, r = threadIdx.x >> 6;
  const int p = blockIdx.x * 64 + l;
  float s = 0.0f;
  if (p <= MLP_NPARAM)
    for (int w = r; w < nwave; w += kReduceSlots) s += part[(size_t)w * (MLP_NPARAM + 1) + p];
  sums[r][l] = s;
  __syncthreads();
  if (r == 0 && p <= MLP_NPARAM) {
    float t = sums[0][l];
    for (int q = 1; q < kReduceSlots; ++q) t += sums[q][l];
    grad[p] = t;
  }
}

// Adam (Kingma & Ba 2015, Algorithm 1) on the FP32 master parameters
__global__ void k_adam(float *__restrict__ p, const float *__restrict__ grad, float *__restrict__ m,
                       float *__restrict__ v, int step, float lr, float beta1, float beta2, float eps) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MLP_NPARAM) return;
  const float gi = grad[i];
  const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
  const float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.0f - powf(beta1, (float)step));
  const float vh = vi / (1.0f - powf(beta2, (float)step));
  p[i] -= lr * mh / (sqrtf(vh) + eps);
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
hipError_t mlp_pack(const float *params, PackedWeights *w, hipStream_t st) {
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_pack, dim3(1), dim3(64), 0, st, params, w);
  return hipGetLastError();
}

hipError_t mlp_forward(const PackedWeights *w, const uint16_t *x, long n, float *y, hipStream_t st) {
  (void)hipGetLastError();
  if (n <= 0) return hipSuccess;
  const long ntile = (n + 15) / 16;
  const long blocks = std::min<long>((ntile + 3) / 4, 4096);
  hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)blocks), dim3(256), 0, st, w, x, n, y);
  return hipGetLastError();
}

int mlp_grad_waves(long n) {
  const long nchunk = (n + 31) / 32;  // one wave per 32 rows
  return (int)((nchunk + 3) / 4 * 4);
}

hipError_t mlp_grad(const PackedWeights *w, const uint16_t *x, const float *t, const uint32_t *idx, long n,
                    float scale, float *part, int nwaves, float *grad, hipStream_t st) {
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_mlp_grad, dim3(nwaves / 4), dim3(256), 0, st, w, x, t, idx, n, scale, part);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAM + 1 + 63) / 64), dim3(64 * kReduceSlots), 0, st, part, nwaves,
                     grad);
  return hipGetLastError();
}

hipError_t mlp_adam(float *p, const float *grad, float *m, float *v, int step, float lr, float beta1, float beta2,
                    float eps, hipStream_t st) {
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, st, p, grad, m, v, step, lr, beta1,
                     beta2, eps);
  return hipGetLastError();
}

}  // namespace f3d
