// Wave-cooperative MLP inference (BF16 MFMA, FP32 accumulation) and the
// surrogate return map used by the PHYS_EP_MLP element kernels.
#pragma once

#include "device.h"
#include "mlp.h"

namespace cpi {

// LDS bytes each wave needs for staging inputs/outputs
constexpr int kMlpStageBytes = 64 * 16 * 2;

// Makes this wave's LDS writes visible to its own later LDS reads.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// All 64 lanes of the wave must call this together.  x: the lane's 13
// standardized inputs; y: its 7 standardized outputs.  stage: this wave's
// kMlpStageBytes of LDS.  Each hidden tile is packed to BF16 as soon as it
// is produced, and the output layer accumulates per K-step, so few hidden
// values are live at once.
__device__ __forceinline__ void mlp_wave_infer(const MlpParams &P, const float x[kMlpIn],
                                               float y[kMlpOut], uint16_t *stage) {
  const int l = threadIdx.x & 63, n = l & 15, g = l >> 4;
  // inputs as bf16 rows [64 samples][16 features]
#pragma unroll
  for (int k = 0; k < 16; k += 2) {
    const uint32_t lo = f2bf(x[k]), hi = k + 1 < kMlpIn ? f2bf(x[k + 1]) : 0u;
    *(uint32_t *)(stage + l * 16 + k) = k < kMlpIn ? lo | (hi << 16) : 0u;
  }
  wave_lds_sync();
  const float zf[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const bf16x8 zero8 = pack_bf8(zf);
  float out[4][4];
#pragma unroll 1
  for (int st = 0; st < 4; st++) {
    // keeps the weight loads inside the loop (hoisted, they would hold ~90
    // registers for the whole element kernel)
    __asm__ volatile("" ::: "memory");
    // layer-1 B operand: features 8g..8g+7 of sample 16st+n (g >= 2: the
    // zero padding of K = 32)
    const bf16x8 xb = g < 2 ? ld_bf8(stage + (16 * st + n) * 16 + 8 * g) : zero8;
    bf16x8 hb[2];
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
      f32x4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = {0.f, 0.f, 0.f, 0.f};
      a0 = mfma16(ld_bf8(P.w1 + (32 * kk + n) * 32 + 8 * g), xb, a0);
      a1 = mfma16(ld_bf8(P.w1 + (32 * kk + 16 + n) * 32 + 8 * g), xb, a1);
      float v[8];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        v[j] = silu(a0[j] + P.b1[32 * kk + 4 * g + j]);
        v[4 + j] = silu(a1[j] + P.b1[32 * kk + 16 + 4 * g + j]);
      }
      hb[kk] = pack_bf8(v);
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
      f32x4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k2 = 0; k2 < 2; k2++) {
        a0 = mfma16(ld_bf8(P.w2p + (32 * kk + n) * 64 + 32 * k2 + 8 * g), hb[k2], a0);
        a1 = mfma16(ld_bf8(P.w2p + (32 * kk + 16 + n) * 64 + 32 * k2 + 8 * g), hb[k2], a1);
      }
      float v[8];
#pragma unroll
      for (int j = 0; j < 4; j++) {
        v[j] = silu(a0[j] + P.b2[32 * kk + 4 * g + j]);
        v[4 + j] = silu(a1[j] + P.b2[32 * kk + 16 + 4 * g + j]);
      }
      // output layer (16 rows, 7 used), K-step kk
      acc = mfma16(ld_bf8(P.w3p + n * 64 + 32 * kk + 8 * g), pack_bf8(v), acc);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) out[st][j] = acc[j] + P.b3[4 * g + j];
  }
  wave_lds_sync();
  // lane n+16g holds outputs 4g..4g+3 of samples 16st+n: through LDS to the
  // sample's own lane
  float *o = (float *)stage;  // [64][8]
  if (g < 2)
#pragma unroll
    for (int st = 0; st < 4; st++)
      *(float4 *)(o + (16 * st + n) * 8 + 4 * g) =
          make_float4(out[st][0], out[st][1], out[st][2], out[st][3]);
  wave_lds_sync();
#pragma unroll
  for (int k = 0; k < kMlpOut; k++) y[k] = o[l * 8 + k];
  wave_lds_sync();
}

// Surrogate return map.  Elastic trial points keep the trial stress exactly
// (no network); if any lane of the wave yields, the wave evaluates the
// network for all its lanes, and each yielding lane accepts the surrogate's
// stress and plastic strain if its yield-function residual is within
// mlp_tol * sy0, else falls back to the exact radial return.  s holds the
// trial stress on entry.  Must be reached by all lanes of the wave together
// (inactive lanes pass active = false).  how: MLP_ELASTIC / MLP_ACCEPTED /
// MLP_FALLBACK.
template <typename R>
__device__ __forceinline__ R mlp_return_how(const Params &p, const R de[6], const R so[6], R &ep,
                                            R s[6], bool active, int &how) {
  const MlpParams &P = *p.mlp;
  const R third = R(1) / R(3);
  const R k23 = R(0.81649658092772603273);
  const R sy0 = R(p.sy0), H = R(p.H);
  R pm = (s[0] + s[1] + s[2]) * third;
  R d0 = s[0] - pm, d1 = s[1] - pm, d2 = s[2] - pm;
  R nrm = sqrt(d0 * d0 + d1 * d1 + d2 * d2 + R(2) * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
  const R q = nrm - k23 * (sy0 + H * ep);  // trial overshoot
  const bool plastic = active && q > R(0);
  how = MLP_ELASTIC;
  const uint64_t mask = __ballot(plastic);
  if (mask == 0) return R(0);
  float x[kMlpIn], y[kMlpOut];
#pragma unroll
  for (int c = 0; c < 6; c++) {
    x[c] = (float(de[c]) - P.in_mean[c]) * P.in_istd[c];
    x[6 + c] = (float(so[c]) - P.in_mean[6 + c]) * P.in_istd[6 + c];
  }
  x[12] = (float(ep) - P.in_mean[12]) * P.in_istd[12];
  extern __shared__ double smem_dyn[];
  uint16_t *stage = (uint16_t *)((char *)smem_dyn + p.mlp_stage_offset) +
                    (threadIdx.x >> 6) * (kMlpStageBytes / 2);
  mlp_wave_infer(P, x, y, stage);
  R dg = R(0);
  if (plastic) {
    R sn[6];
#pragma unroll
    for (int c = 0; c < 6; c++) sn[c] = s[c] - R(y[c] * P.out_std[c] + P.out_mean[c]) * q;
    R dep = R(y[6] * P.out_std[6] + P.out_mean[6]) * q / sy0;
    dep = dep > R(0) ? dep : R(0);
    const R epn = ep + dep;
    R qm = (sn[0] + sn[1] + sn[2]) * third;
    R e0 = sn[0] - qm, e1 = sn[1] - qm, e2 = sn[2] - qm;
    R qn = sqrt(e0 * e0 + e1 * e1 + e2 * e2 + R(2) * (sn[3] * sn[3] + sn[4] * sn[4] + sn[5] * sn[5]));
    R res = qn - k23 * (sy0 + H * epn);
    if (fabs(res) <= R(p.mlp_tol) * sy0) {
#pragma unroll
      for (int c = 0; c < 6; c++) s[c] = sn[c];
      dg = dep / k23;
      ep = epn;
      how = MLP_ACCEPTED;
    } else {
      dg = radial_return<R>(p, s, ep);
      how = MLP_FALLBACK;
    }
  }
  if (p.mlp_stats) {
    const uint64_t fb = __ballot(how == MLP_FALLBACK);
    if ((threadIdx.x & 63) == 0) {
      atomicAdd(p.mlp_stats, (unsigned long long)(__popcll(mask) - __popcll(fb)));
      atomicAdd(p.mlp_stats + 1, (unsigned long long)__popcll(fb));
    }
  }
  return dg;
}

template <typename R>
__device__ R mlp_return(const Params &p, const R de[6], const R so[6], R &ep, R s[6],
                        bool active) {
  int how;
  return mlp_return_how<R>(p, de, so, ep, s, active, how);
}

}  // namespace cpi
