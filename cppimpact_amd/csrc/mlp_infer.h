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

// All 64 lanes of the wave must call this together.  Lanes with want set
// (mask = the wave's ballot of want) pass their 13 standardized inputs in x
// and get the 7 standardized outputs in y.  Their samples are packed into
// the first ceil(popc(mask)/16) 16-sample tiles, so the matrix cores work
// only on points that need the network.  stage: this wave's kMlpStageBytes
// of LDS; sample row r holds 16 bf16 inputs, later 8 float outputs (both 32
// bytes, so a tile's outputs overwrite its own inputs only).  Each hidden
// tile is packed to BF16 as soon as it is produced and the output layer
// accumulates per K-step, so few hidden values are live at once.
__device__ __forceinline__ void mlp_wave_infer(const MlpParams &P, const float x[kMlpIn],
                                               float y[kMlpOut], uint16_t *stage, bool want,
                                               uint64_t mask) {
  const int l = threadIdx.x & 63, n = l & 15, g = l >> 4;
  const int rank = __popcll(mask & ((1ull << l) - 1ull));
  const int ntiles = (__popcll(mask) + 15) >> 4;
  if (want) {
#pragma unroll
    for (int k = 0; k < 16; k += 2) {
      const uint32_t lo = f2bf(x[k]), hi = k + 1 < kMlpIn ? f2bf(x[k + 1]) : 0u;
      *(uint32_t *)(stage + rank * 16 + k) = k < kMlpIn ? lo | (hi << 16) : 0u;
    }
  }
  wave_lds_sync();
  const float zf[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const bf16x8 zero8 = pack_bf8(zf);
  float *o = (float *)stage;  // [64][8] outputs
#pragma unroll 1
  for (int st = 0; st < ntiles; st++) {
    // keeps the weight loads inside the loop (hoisted, they would hold ~90
    // registers for the whole element kernel)
    __asm__ volatile("" ::: "memory");
    // layer-1 B operand: features 8g..8g+7 of sample 16st+n (g >= 2: the
    // zero padding of K = 32).  Rows past the packed samples hold stale
    // data; their output columns are computed and never read.
    const bf16x8 xb = g < 2 ? ld_bf8(stage + (16 * st + n) * 16 + 8 * g) : zero8;
    // the scheduling barriers keep each group's weight and bias loads next
    // to their MFMAs: hoisted together they need ~100 registers
    bf16x8 hb[2];
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
      __builtin_amdgcn_sched_barrier(0);
      f32x4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = {0.f, 0.f, 0.f, 0.f};
      a0 = mfma16(ld_bf8(P.w1 + (32 * kk + n) * 32 + 8 * g), xb, a0);
      a1 = mfma16(ld_bf8(P.w1 + (32 * kk + 16 + n) * 32 + 8 * g), xb, a1);
      const float4 c0 = *(const float4 *)(P.b1 + 32 * kk + 4 * g);
      const float4 c1 = *(const float4 *)(P.b1 + 32 * kk + 16 + 4 * g);
      const float v[8] = {silu(a0[0] + c0.x), silu(a0[1] + c0.y), silu(a0[2] + c0.z),
                          silu(a0[3] + c0.w), silu(a1[0] + c1.x), silu(a1[1] + c1.y),
                          silu(a1[2] + c1.z), silu(a1[3] + c1.w)};
      hb[kk] = pack_bf8(v);
    }
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
      f32x4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int k2 = 0; k2 < 2; k2++) {
        __builtin_amdgcn_sched_barrier(0);
        a0 = mfma16(ld_bf8(P.w2p + (32 * kk + n) * 64 + 32 * k2 + 8 * g), hb[k2], a0);
        a1 = mfma16(ld_bf8(P.w2p + (32 * kk + 16 + n) * 64 + 32 * k2 + 8 * g), hb[k2], a1);
      }
      __builtin_amdgcn_sched_barrier(0);
      const float4 c0 = *(const float4 *)(P.b2 + 32 * kk + 4 * g);
      const float4 c1 = *(const float4 *)(P.b2 + 32 * kk + 16 + 4 * g);
      const float v[8] = {silu(a0[0] + c0.x), silu(a0[1] + c0.y), silu(a0[2] + c0.z),
                          silu(a0[3] + c0.w), silu(a1[0] + c1.x), silu(a1[1] + c1.y),
                          silu(a1[2] + c1.z), silu(a1[3] + c1.w)};
      // output layer (16 rows, 7 used), K-step kk
      acc = mfma16(ld_bf8(P.w3p + n * 64 + 32 * kk + 8 * g), pack_bf8(v), acc);
    }
    __builtin_amdgcn_sched_barrier(0);
    // lane n+16g holds outputs 4g..4g+3 of sample 16st+n; the lane's input
    // read of that row (xb) has completed, as the MFMAs above consumed it
    if (g < 2)
      *(float4 *)(o + (16 * st + n) * 8 + 4 * g) =
          make_float4(acc[0] + P.b3[4 * g], acc[1] + P.b3[4 * g + 1], acc[2] + P.b3[4 * g + 2],
                      acc[3] + P.b3[4 * g + 3]);
  }
  wave_lds_sync();
  if (want)
#pragma unroll
    for (int k = 0; k < kMlpOut; k++) y[k] = o[rank * 8 + k];
  wave_lds_sync();
}

// Standardized network inputs of one point: (de, sigma_n, ep_n).
template <typename R>
__device__ __forceinline__ void mlp_inputs(const MlpParams &P, const R de[6], const R so[6], R ep,
                                           float x[kMlpIn]) {
#pragma unroll
  for (int c = 0; c < 6; c++) {
    x[c] = (float(de[c]) - P.in_mean[c]) * P.in_istd[c];
    x[6 + c] = (float(so[c]) - P.in_mean[6 + c]) * P.in_istd[6 + c];
  }
  x[12] = (float(ep) - P.in_mean[12]) * P.in_istd[12];
}

// The surrogate's return for one yielding point (q > 0): s holds the trial
// stress.  The network's outputs y give sigma_{n+1} = s - y_c q and
// ep_{n+1} = ep + max(y_6 q / sy0, 0); they are accepted if the yield
// residual is within mlp_tol * sy0, else the exact radial return is taken.
// Returns the plastic multiplier; how = MLP_ACCEPTED or MLP_FALLBACK.
template <typename R>
__device__ __forceinline__ R mlp_apply(const Params &p, const float y[kMlpOut], R q, R &ep,
                                       R s[6], int &how) {
  const MlpParams &P = *p.mlp;
  const R third = R(1) / R(3);
  const R k23 = R(0.81649658092772603273);
  const R sy0 = R(p.sy0), H = R(p.H);
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
    ep = epn;
    how = MLP_ACCEPTED;
    return dep / k23;
  }
  how = MLP_FALLBACK;
  return radial_return<R>(p, s, ep);
}

// accepted / fell-back counts of one wave, added to the workgroup's pair in
// LDS (the workgroup adds the pair to its own global counters once)
__device__ __forceinline__ void mlp_count(unsigned *cnt, bool accepted, bool fell_back) {
  const uint64_t a = __ballot(accepted), f = __ballot(fell_back);
  if ((threadIdx.x & 63) == 0 && (a | f)) {
    atomicAdd(cnt, (unsigned)__popcll(a));
    atomicAdd(cnt + 1, (unsigned)__popcll(f));
  }
}

// this wave's MLP staging area in the dynamic LDS
__device__ __forceinline__ uint16_t *mlp_stage(const Params &p) {
  extern __shared__ double smem_dyn[];
  return (uint16_t *)((char *)smem_dyn + p.mlp_stage_offset) +
         (threadIdx.x >> 6) * (kMlpStageBytes / 2);
}

}  // namespace cpi
