// MLP surrogate of the J2 return map: shared device definitions.
//
// Network 13 -> 64 -> 64 -> 7, SiLU on the hidden layers.  Inputs are the
// standardized (strain increment[6], stress_n[6], eq. plastic strain_n).
// Outputs are standardized and scaled by the trial overshoot
// q = |dev sigma_trial| - sqrt(2/3) (sy0 + H ep_n) > 0: the return-map
// correction (sigma_trial - sigma_{n+1}) / q [6] and sy0 * dep / q, so the
// surrogate's error shrinks with the overshoot it corrects.  Weights are
// BF16 on the MFMA units (v_mfma_f32_16x16x32_bf16), accumulation FP32.
//
// Every matrix product is computed transposed, H^T = W X^T, so a layer's
// output tile (rows = units, columns = samples) is already laid out as the
// next layer's B operand: in a 16x16x32 MFMA, lane i+16g of A and lane
// n+16g of B each supply 8 values of the K index, and the products pair
// element j of A-lane (i,g) with element j of B-lane (n,g).  Taking the two
// 16-unit output tiles 2kk and 2kk+1 as K-step kk, B-lane (n,g) holds units
// 32kk+4g+j (j<4) and 32kk+16+4g+(j-4) (j>=4); the weights of the next layer
// are stored permuted the same way (W_perm), so no transpose is needed.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device.h"

namespace cpi {

constexpr int kMlpIn = 13, kMlpHid = 64, kMlpOut = 7;
constexpr int kSampleFloats = kMlpIn + kMlpOut;
// parameters in canonical order: W1[64][13], b1[64], W2[64][64], b2[64],
// W3[7][64], b3[7]
constexpr int kMlpW1 = 0, kMlpB1 = kMlpW1 + 64 * 13, kMlpW2 = kMlpB1 + 64,
              kMlpB2 = kMlpW2 + 64 * 64, kMlpW3 = kMlpB2 + 64, kMlpB3 = kMlpW3 + 7 * 64,
              kMlpNParams = kMlpB3 + 7;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// Device image of the network as the inference kernels read it (16-byte
// aligned arrays).
struct MlpWeights {
  uint16_t w1[64 * 32];   // W1[u][k], k >= 13 zero
  uint16_t w2p[64 * 64];  // W2, K permuted (perm_k)
  uint16_t w3p[16 * 64];  // W3, rows >= 7 zero, K permuted
  float b1[64], b2[64], b3[16];
};

// Extra layouts of the same weights for the training kernels.
struct MlpTrainWeights {
  uint16_t w2[64 * 64];   // W2[u][v]
  uint16_t w2t[64 * 64];  // W2^T[v][u]
  uint16_t w3[16 * 64];   // W3[o][u], rows >= 7 zero
  uint16_t w3t[64 * 32];  // W3^T[u][o], o >= 7 zero
};

// Weights as the kernels read them.
struct MlpParams {
  const uint16_t *w1;   // [64][32]  W1[u][k], k >= 13 zero (K = input feature)
  const uint16_t *w2p;  // [64][64]  W2 with K permuted as described above
  const uint16_t *w3p;  // [16][64]  W3 (rows >= 7 zero), K permuted
  const float *b1, *b2, *b3;  // [64] [64] [16]
  float in_mean[kMlpIn], in_istd[kMlpIn];
  float out_mean[kMlpOut], out_std[kMlpOut];
  // material of the return map the network stands for
  float c11, c12, c44, sy0, H;
};

// round-to-nearest-even float -> bf16 bits
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 ld_bf8(const uint16_t *p) { return *(const bf16x8 *)p; }

__device__ __forceinline__ bf16x8 pack_bf8(const float v[8]) {
  union {
    uint16_t h[8];
    bf16x8 b;
  } u;
#pragma unroll
  for (int j = 0; j < 8; j++) u.h[j] = f2bf(v[j]);
  return u.b;
}

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// C tile (16x16) = sum over k in [0, K) of A[i][k] * Bt[n][k], with A and Bt
// row-major bf16 (K contiguous, leading dimensions lda/ldb, rows 16-byte
// aligned).  Lane n+16g receives C[4g+j][n], j = 0..3.
__device__ __forceinline__ f32x4 mfma_tile(const uint16_t *A, int lda, const uint16_t *Bt, int ldb,
                                          int K, f32x4 acc) {
  const int l = threadIdx.x & 63, i = l & 15, g = l >> 4;
  for (int k0 = 0; k0 < K; k0 += 32)
    acc = mfma16(ld_bf8(A + i * lda + k0 + 8 * g), ld_bf8(Bt + i * ldb + k0 + 8 * g), acc);
  return acc;
}

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// index of K-permuted weight storage: unit k of K-step kk sits at
// kk*32 + g*8 + j with j<4: k = 32kk+4g+j, j>=4: k = 32kk+16+4g+(j-4)
__host__ __device__ __forceinline__ int perm_k(int k) {
  const int kk = k >> 5, r = k & 31, hi = r >> 4, g = (r & 15) >> 2, j = r & 3;
  return kk * 32 + g * 8 + hi * 4 + j;
}

// Records (de, sigma_n, ep_n) -> (sigma_{n+1}, ep_{n+1}) for points that
// yielded.  Slots come from an atomic counter; samples past the capacity are
// dropped (the counter still counts them).
template <typename R>
__device__ __forceinline__ void record_sample(const Params &p, const R de[6], const R so[6],
                                              R ep0, const R s[6], R ep1) {
  int slot = atomicAdd(p.sample_count, 1);
  if (slot >= p.sample_cap) return;
  float *d = p.samples + (long)slot * kSampleFloats;
#pragma unroll
  for (int c = 0; c < 6; c++) {
    d[c] = float(de[c]);
    d[6 + c] = float(so[c]);
    d[13 + c] = float(s[c]);
  }
  d[12] = float(ep0);
  d[19] = float(ep1);
}

// outcome of one point of the surrogate return map
enum : int { MLP_ELASTIC = 0, MLP_ACCEPTED = 1, MLP_FALLBACK = 2 };

}  // namespace cpi
