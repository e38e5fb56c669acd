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
  const MlpWeights *weights;  // the whole image (copied into LDS per workgroup)
  const uint16_t *w1;   // [64][32]  W1[u][k], k >= 13 zero (K = input feature)
  const uint16_t *w2p;  // [64][64]  W2 with K permuted as described above
  const uint16_t *w3p;  // [16][64]  W3 (rows >= 7 zero), K permuted
  const float *b1, *b2, *b3;  // [64] [64] [16]
  float in_mean[kMlpIn], in_istd[kMlpIn];
  float out_mean[kMlpOut], out_std[kMlpOut];
  // material of the return map the network stands for
  float c11, c12, c44, sy0, H;
};

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// two floats -> packed bf16 bits (a low), round to nearest even: one
// v_cvt_pk_bf16_f32 on gfx950 instead of a handful of integer operations
// per value
__device__ __forceinline__ uint32_t pack_bf2(float a, float b) {
  const f32x2 v = {a, b};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// round-to-nearest-even float -> bf16 bits
__device__ __forceinline__ uint16_t f2bf(float f) { return (uint16_t)pack_bf2(f, 0.f); }
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 ld_bf8(const uint16_t *p) { return *(const bf16x8 *)p; }

// Makes this wave's LDS writes visible to its own later LDS reads.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bf16x8 pack_bf8(const float v[8]) {
  union {
    uint32_t w[4];
    bf16x8 b;
  } u;
#pragma unroll
  for (int j = 0; j < 4; j++) u.w[j] = pack_bf2(v[2 * j], v[2 * j + 1]);
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

// z sigmoid(z) with the hardware exponential and reciprocal (1 ulp each): a
// correctly rounded division would cost a dozen instructions per unit, and
// the result is rounded to BF16 next anyway
__device__ __forceinline__ float sigmoid(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __expf(-z));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = sigmoid(z);
  return s * (1.0f + z * (1.0f - s));
}

// index of K-permuted weight storage: unit k of K-step kk sits at
// kk*32 + g*8 + j with j<4: k = 32kk+4g+j, j>=4: k = 32kk+16+4g+(j-4)
__host__ __device__ __forceinline__ int perm_k(int k) {
  const int kk = k >> 5, r = k & 31, hi = r >> 4, g = (r & 15) >> 2, j = r & 3;
  return kk * 32 + g * 8 + hi * 4 + j;
}

// SplitMix64 finaliser: the recorder's subsampling hash
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Records (de, sigma_n, ep_n) -> (sigma_{n+1}, ep_{n+1}) for points that
// yielded, point q of element e in this step, when the hash of (seed, step,
// e, q) falls below the keep threshold (a fixed, reproducible subset).
// Slots come from an atomic counter; samples past the capacity are dropped
// (the counter still counts them).
template <typename R>
__device__ __forceinline__ void record_sample(const Params &p, int e, int q, const R de[6],
                                              const R so[6], R ep0, const R s[6], R ep1) {
  if (p.sample_keep < (1ull << 32)) {
    const uint64_t step = (uint64_t)p.ctl->step;
    const uint64_t h = mix64(p.sample_seed + 0x9E3779B97F4A7C15ull * (step * 8 + q + 1) +
                             (uint64_t)(uint32_t)e * 0xD1B54A32D192ED03ull);
    if ((h >> 32) >= p.sample_keep) return;
  }
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
