// Return-mapping sample recording and the MLP surrogate hook used by the
// element kernels.  The network itself (BF16 MFMA forward/backward) lives in
// mlp.hip; this header holds what the element kernels need.
#pragma once

#include "device.h"

namespace cpi {

constexpr int kMlpIn = 13, kMlpHid = 64, kMlpOut = 7;
constexpr int kSampleFloats = kMlpIn + kMlpOut;

// Weights as the inference kernels read them: bf16 matrices padded to MFMA
// tiles, f32 biases, and the input/output scalings.
struct MlpParams {
  const uint16_t *w1;  // [64][32]  (in 13 padded to 32), row = output unit
  const uint16_t *w2;  // [64][64]
  const uint16_t *w3;  // [16][64]  (out 7 padded to 16)
  const float *b1, *b2, *b3;
  float in_scale[kMlpIn];
  float out_scale[kMlpOut];
};

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

// Surrogate return map; defined in mlp_infer.h (wave-cooperative MFMA).
template <typename R>
__device__ R mlp_return(const Params &p, const R de[6], const R so[6], R &ep, R s[6]);

}  // namespace cpi
