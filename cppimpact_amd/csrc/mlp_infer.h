// MLP inference for the surrogate return map (BF16 MFMA, FP32
// accumulation) and the return map built on it.
//
// A surrogate step is three launches (kernels.hip): the elastic pass writes
// the standardized inputs of every yielding point of a block it defers to a
// dense queue of 32-byte rows (mlp_store_inputs); k_mlp_queue evaluates the
// network on the whole queue, 16 samples per MFMA tile, with the weights in
// LDS (mlp_tile), writing each row's outputs over its inputs; the deferred
// blocks are then finished with the outputs (mlp_apply).
#pragma once

#include "device.h"
#include "mlp.h"

namespace cpi {

// bytes of one queue row: 16 BF16 inputs (13 used), later 8 FP32 outputs
// (7 used)
constexpr int kMlpRowBytes = 32;
constexpr int kMlpRowHalves = kMlpRowBytes / 2;

// The workgroup's LDS copy of the weights, in the layout of MlpWeights.
__device__ __forceinline__ void mlp_lds_load(const MlpWeights *src, MlpWeights *dst) {
  static_assert(sizeof(MlpWeights) % 16 == 0, "weights copied as 16-byte words");
  const uint4 *s = reinterpret_cast<const uint4 *>(src);
  uint4 *d = reinterpret_cast<uint4 *>(dst);
  for (int i = threadIdx.x; i < (int)(sizeof(MlpWeights) / 16); i += blockDim.x) d[i] = s[i];
}

// The layer-1 B operand of one tile of 16 queue rows for this lane: lane
// n + 16g (g < 2) takes features 8g..8g+7 of row n, lanes with g >= 2 the
// zero padding of K = 32.
__device__ __forceinline__ bf16x8 mlp_tile_inputs(const uint16_t *rows) {
  const int l = threadIdx.x & 63, n = l & 15, g = l >> 4;
  const float zf[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  return g < 2 ? ld_bf8(rows + n * kMlpRowHalves + 8 * g) : pack_bf8(zf);
}

// The network on one tile of 16 queue rows (rows[16 * kMlpRowHalves]),
// whose inputs xb were read by mlp_tile_inputs: row n holds sample n's
// standardized BF16 inputs and receives its 7 FP32 outputs.  All 64 lanes
// of the wave take part.  Every product is computed transposed (H^T = W
// X^T, see mlp.h), so lane n + 16g holds units 4g..4g+3 of each 16-unit
// output tile of sample n, already the layout of the next layer's B
// operand.  Row n is written by lanes n and n + 16, the lanes that read it.
__device__ __forceinline__ void mlp_tile(const MlpWeights &W, const bf16x8 xb, uint16_t *rows) {
  const int l = threadIdx.x & 63, n = l & 15, g = l >> 4;
  // the scheduling barriers keep each group's LDS weight reads next to
  // their MFMAs: hoisted together the 14 weight fragments hold 56 VGPRs
  bf16x8 hb[2];
#pragma unroll
  for (int kk = 0; kk < 2; kk++) {
    __builtin_amdgcn_sched_barrier(0);
    f32x4 a0 = {0.f, 0.f, 0.f, 0.f}, a1 = {0.f, 0.f, 0.f, 0.f};
    a0 = mfma16(ld_bf8(W.w1 + (32 * kk + n) * 32 + 8 * g), xb, a0);
    a1 = mfma16(ld_bf8(W.w1 + (32 * kk + 16 + n) * 32 + 8 * g), xb, a1);
    const float4 c0 = *(const float4 *)(W.b1 + 32 * kk + 4 * g);
    const float4 c1 = *(const float4 *)(W.b1 + 32 * kk + 16 + 4 * g);
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
      a0 = mfma16(ld_bf8(W.w2p + (32 * kk + n) * 64 + 32 * k2 + 8 * g), hb[k2], a0);
      a1 = mfma16(ld_bf8(W.w2p + (32 * kk + 16 + n) * 64 + 32 * k2 + 8 * g), hb[k2], a1);
    }
    __builtin_amdgcn_sched_barrier(0);
    const float4 c0 = *(const float4 *)(W.b2 + 32 * kk + 4 * g);
    const float4 c1 = *(const float4 *)(W.b2 + 32 * kk + 16 + 4 * g);
    const float v[8] = {silu(a0[0] + c0.x), silu(a0[1] + c0.y), silu(a0[2] + c0.z),
                        silu(a0[3] + c0.w), silu(a1[0] + c1.x), silu(a1[1] + c1.y),
                        silu(a1[2] + c1.z), silu(a1[3] + c1.w)};
    // output layer (16 rows, 7 used), K-step kk
    acc = mfma16(ld_bf8(W.w3p + n * 64 + 32 * kk + 8 * g), pack_bf8(v), acc);
  }
  if (g < 2) {
    const float4 c = *(const float4 *)(W.b3 + 4 * g);
    *(float4 *)((float *)rows + n * (kMlpRowBytes / 4) + 4 * g) =
        make_float4(acc[0] + c.x, acc[1] + c.y, acc[2] + c.z, acc[3] + c.w);
  }
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

// One queue row: the 13 inputs as BF16 (round to nearest even), zero padded.
__device__ __forceinline__ void mlp_store_inputs(uint16_t *row, const float x[kMlpIn]) {
  uint32_t w[8];
#pragma unroll
  for (int k = 0; k < 8; k++)
    w[k] = pack_bf2(2 * k < kMlpIn ? x[2 * k] : 0.f, 2 * k + 1 < kMlpIn ? x[2 * k + 1] : 0.f);
  uint4 *d = reinterpret_cast<uint4 *>(row);
  d[0] = make_uint4(w[0], w[1], w[2], w[3]);
  d[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// The 7 outputs of one queue row after k_mlp_queue.
__device__ __forceinline__ void mlp_load_outputs(const uint16_t *row, float y[kMlpOut]) {
  const float4 *s = reinterpret_cast<const float4 *>(row);
  const float4 a = s[0], b = s[1];
  y[0] = a.x;
  y[1] = a.y;
  y[2] = a.z;
  y[3] = a.w;
  y[4] = b.x;
  y[5] = b.y;
  y[6] = b.z;
}

// The exact radial return in FP64 on an FP32 trial state, rounded back:
// the surrogate's fallback.
template <typename R>
__device__ __forceinline__ R radial_return_f64(const Params &p, R s[6], R &ep) {
  double sd[6], ed = double(ep);
#pragma unroll
  for (int c = 0; c < 6; c++) sd[c] = double(s[c]);
  const double dg = radial_return<double>(p, sd, ed);
#pragma unroll
  for (int c = 0; c < 6; c++) s[c] = R(sd[c]);
  ep = R(ed);
  return R(dg);
}

// The surrogate's return for one yielding point (q > 0): s holds the trial
// stress.  The network's outputs y give the correction c = y_c q, so
// sigma_{n+1} = s - c, and ep_{n+1} = ep + max(y_6 q / sy0, 0).  A J2
// return keeps the pressure, scales the trial deviator along itself, lowers
// its norm by sqrt(6) G dep and ends on the yield surface; the network's
// state is accepted only if it does all four within mlp_tol * sy0:
//   |tr c| / 3                                        (pressure kept)
//   |dev c - (dev c : n) n|, n = dev s / |dev s|      (no turn)
//   ||dev s| - |dev s_{n+1}| - sqrt(6) G dep|         (consistency)
//   ||dev s_{n+1}| - sqrt(2/3) (sy0 + H ep_{n+1})|    (yield residual)
// with |a| = sqrt(a : a) in Voigt form (shear terms twice).  The checks
// work on c, of the size of the overshoot, so FP32 resolves the turn far
// below the tolerance.  Otherwise the exact radial return is taken in
// FP64.  Returns the plastic multiplier; how = MLP_ACCEPTED or
// MLP_FALLBACK.
template <typename R>
__device__ __forceinline__ R mlp_apply(const Params &p, const float y[kMlpOut], R q, R &ep,
                                       R s[6], int &how) {
  const MlpParams &P = *p.mlp;
  const R third = R(1) / R(3);
  const R k23 = R(0.81649658092772603273);  // sqrt(2/3)
  const R r6 = R(2.44948974278317809820);   // sqrt(6)
  const R sy0 = R(p.sy0), H = R(p.H), G = R(p.mu);
  R c[6];
#pragma unroll
  for (int k = 0; k < 6; k++) c[k] = R(y[k] * P.out_std[k] + P.out_mean[k]) * q;
  R dep = R(y[6] * P.out_std[6] + P.out_mean[6]) * q / sy0;
  dep = dep > R(0) ? dep : R(0);
  const R pt = (s[0] + s[1] + s[2]) * third, pc = (c[0] + c[1] + c[2]) * third;
  R tt = R(0), ct = R(0), cc = R(0);
#pragma unroll
  for (int k = 0; k < 6; k++) {
    const R w = k < 3 ? R(1) : R(2);
    const R dt = k < 3 ? s[k] - pt : s[k], dc = k < 3 ? c[k] - pc : c[k];
    tt += w * dt * dt;
    ct += w * dc * dt;
    cc += w * dc * dc;
  }
  const R nt = sqrt(tt), nn = sqrt(tt - R(2) * ct + cc);
  const R perp2 = cc - ct * ct / tt;
  const R tol = R(p.mlp_tol) * sy0;
  const bool ok = fabs(pc) <= tol && perp2 <= tol * tol && fabs((nt - nn) - r6 * G * dep) <= tol &&
                  fabs(nn - k23 * (sy0 + H * (ep + dep))) <= tol;
  if (ok) {
#pragma unroll
    for (int k = 0; k < 6; k++) s[k] -= c[k];
    ep += dep;
    how = MLP_ACCEPTED;
    return dep / k23;
  }
  how = MLP_FALLBACK;
  return radial_return_f64<R>(p, s, ep);
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

}  // namespace cpi
