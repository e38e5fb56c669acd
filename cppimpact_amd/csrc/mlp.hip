// MLP surrogate of the J2 return map: training (BF16 MFMA forward and
// backward passes, FP32 accumulation, Adam on FP32 master weights), the
// evaluation kernels behind cpi_mlp_forward / cpi_mlp_return_map, and the
// cpi_mlp C ABI.
//
// Training minimizes L = 1/(2B) sum_s sum_o (y_so - t_so)^2 over standardized
// outputs.  Operands are rounded to BF16 where they enter an MFMA: the
// inputs, the weights, the hidden activations, dL/dy and the pre-activation
// gradients; SiLU derivatives, bias gradients and all sums stay FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/cppimpact_hip.h"
#include "elem.h"
#include "kernels.h"
#include "mlp_host.h"
#include "mlp_infer.h"

namespace cpi {

// 64-sample chunks per gradient launch (mini-batches up to 65536 samples)
constexpr int kMaxParts = 1024;

// ---------------------------------------------------------------------------
// device kernels
// ---------------------------------------------------------------------------

// parameter i (canonical order) into every BF16 layout / bias slot
__device__ __forceinline__ void scatter_param(int i, float v, MlpWeights *W, MlpTrainWeights *T) {
  const uint16_t h = f2bf(v);
  if (i < kMlpB1) {
    W->w1[(i / kMlpIn) * 32 + i % kMlpIn] = h;
  } else if (i < kMlpW2) {
    W->b1[i - kMlpB1] = v;
  } else if (i < kMlpB2) {
    const int r = i - kMlpW2, u = r >> 6, k = r & 63;
    T->w2[u * 64 + k] = h;
    T->w2t[k * 64 + u] = h;
    W->w2p[u * 64 + perm_k(k)] = h;
  } else if (i < kMlpW3) {
    W->b2[i - kMlpB2] = v;
  } else if (i < kMlpB3) {
    const int r = i - kMlpW3, o = r >> 6, u = r & 63;
    T->w3[o * 64 + u] = h;
    T->w3t[u * 32 + o] = h;
    W->w3p[o * 64 + perm_k(u)] = h;
  } else {
    W->b3[i - kMlpB3] = v;
  }
}

__global__ void __launch_bounds__(256) k_mlp_scatter(const float *theta, MlpWeights *W,
                                                     MlpTrainWeights *T) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < kMlpNParams) scatter_param(i, theta[i], W, T);
}

__device__ __forceinline__ float sum16(float v) {
  v += __shfl_xor(v, 1);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 4);
  v += __shfl_xor(v, 8);
  return v;
}

// LDS of one gradient workgroup (one wave, 64 samples); 52 KiB
struct GradLds {
  uint16_t xs[64 * 32];    // X [s][k], k >= 13 zero
  uint16_t xt[16 * 64];    // X^T [k][s]
  uint16_t a1[64 * 64];    // A1 [s][u]; dZ2 [s][u] in the backward pass
  uint16_t a1t[64 * 64];   // A1^T [u][s]
  uint16_t a2[64 * 64];    // A2 [s][u]; dZ1^T [u][s] in the backward pass
  uint16_t a2t[64 * 64];   // A2^T [u][s]
  uint16_t dy[64 * 32];    // dL/dY [s][o], o >= 16 zero
  uint16_t dyt[16 * 64];   // dL/dY^T [o][s]
  uint16_t dz2t[64 * 64];  // dZ2^T [u][s]
};

// Gradient of one mini-batch chunk: workgroup b takes samples
// b0 + 64 b + (0..63) of the order idx (or the identity), and writes its
// parameter gradient (canonical order) to gpart[b] and its loss to lpart[b].
__global__ void __launch_bounds__(64) k_mlp_grad(const float *X, const float *T, const int *idx,
                                                 int b0, int bn, float inv_b, const MlpWeights *W,
                                                 const MlpTrainWeights *TW, float *gpart,
                                                 float *lpart) {
  __shared__ __attribute__((aligned(16))) GradLds L;
  const int l = threadIdx.x, n = l & 15, g = l >> 4;
  const int c0 = blockIdx.x * 64;
  float *gp = gpart + (long)blockIdx.x * kMlpNParams;
  const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
  {
    const int sl = c0 + l;
    const bool ok = sl < bn;
    const long gs = ok ? (idx ? idx[b0 + sl] : b0 + sl) : 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const uint16_t h = f2bf(ok && k < kMlpIn ? X[gs * 16 + k] : 0.f);
      L.xs[l * 32 + k] = h;
      L.xt[k * 64 + l] = h;
    }
#pragma unroll
    for (int k = 16; k < 32; k++) {
      L.xs[l * 32 + k] = 0;
      L.dy[l * 32 + k] = 0;
    }
  }
  wave_lds_sync();
  // SiLU derivatives of both hidden layers, tile (unit tile, sample tile)
  float d1[4][4][4], d2[4][4][4];
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int st = 0; st < 4; st++) {
      const f32x4 acc = mfma_tile(W->w1 + 16 * mt * 32, 32, L.xs + 16 * st * 32, 32, 32, zero);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int u = 16 * mt + 4 * g + j, s = 16 * st + n;
        const float z = acc[j] + W->b1[u];
        d1[mt][st][j] = silu_grad(z);
        const uint16_t h = f2bf(silu(z));
        L.a1[s * 64 + u] = h;
        L.a1t[u * 64 + s] = h;
      }
    }
  wave_lds_sync();
#pragma unroll
  for (int mt = 0; mt < 4; mt++)
#pragma unroll
    for (int st = 0; st < 4; st++) {
      const f32x4 acc = mfma_tile(TW->w2 + 16 * mt * 64, 64, L.a1 + 16 * st * 64, 64, 64, zero);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int u = 16 * mt + 4 * g + j, s = 16 * st + n;
        const float z = acc[j] + W->b2[u];
        d2[mt][st][j] = silu_grad(z);
        const uint16_t h = f2bf(silu(z));
        L.a2[s * 64 + u] = h;
        L.a2t[u * 64 + s] = h;
      }
    }
  wave_lds_sync();
  // output layer, loss and dL/dy
  float loss = 0.f, db3[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int st = 0; st < 4; st++) {
    const f32x4 acc = mfma_tile(TW->w3, 64, L.a2 + 16 * st * 64, 64, 64, zero);
    const int s = 16 * st + n, sl = c0 + s;
    const bool ok = sl < bn;
    const long gs = ok ? (idx ? idx[b0 + sl] : b0 + sl) : 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int o = 4 * g + j;
      float dy = 0.f;
      if (ok && o < kMlpOut) {
        const float r = acc[j] + W->b3[o] - T[gs * 8 + o];
        loss += 0.5f * r * r * inv_b;
        dy = r * inv_b;
      }
      db3[j] += dy;
      const uint16_t h = f2bf(dy);
      L.dy[s * 32 + o] = h;
      L.dyt[o * 64 + s] = h;
    }
  }
  wave_lds_sync();
  // dW3[o][u] = dY^T A2
#pragma unroll
  for (int ut = 0; ut < 4; ut++) {
    const f32x4 acc = mfma_tile(L.dyt, 64, L.a2t + 16 * ut * 64, 64, 64, zero);
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (4 * g + j < kMlpOut) gp[kMlpW3 + (4 * g + j) * 64 + 16 * ut + n] = acc[j];
  }
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const float v = sum16(db3[j]);
    if (n == 0 && 4 * g + j < kMlpOut) gp[kMlpB3 + 4 * g + j] = v;
  }
  // dZ2^T = (W3^T dY^T) * silu'(Z2)
  float db2[4][4];
#pragma unroll
  for (int ut = 0; ut < 4; ut++) {
#pragma unroll
    for (int j = 0; j < 4; j++) db2[ut][j] = 0.f;
#pragma unroll
    for (int st = 0; st < 4; st++) {
      const f32x4 acc = mfma_tile(TW->w3t + 16 * ut * 32, 32, L.dy + 16 * st * 32, 32, 32, zero);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int u = 16 * ut + 4 * g + j, s = 16 * st + n;
        const float dz = acc[j] * d2[ut][st][j];
        db2[ut][j] += dz;
        const uint16_t h = f2bf(dz);
        L.a1[s * 64 + u] = h;
        L.dz2t[u * 64 + s] = h;
      }
    }
  }
  wave_lds_sync();
  // dW2[u][v] = dZ2^T A1
#pragma unroll
  for (int ut = 0; ut < 4; ut++) {
#pragma unroll
    for (int vt = 0; vt < 4; vt++) {
      const f32x4 acc = mfma_tile(L.dz2t + 16 * ut * 64, 64, L.a1t + 16 * vt * 64, 64, 64, zero);
#pragma unroll
      for (int j = 0; j < 4; j++) gp[kMlpW2 + (16 * ut + 4 * g + j) * 64 + 16 * vt + n] = acc[j];
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const float v = sum16(db2[ut][j]);
      if (n == 0) gp[kMlpB2 + 16 * ut + 4 * g + j] = v;
    }
  }
  // dZ1^T = (W2^T dZ2^T) * silu'(Z1)
  float db1[4][4];
#pragma unroll
  for (int vt = 0; vt < 4; vt++) {
#pragma unroll
    for (int j = 0; j < 4; j++) db1[vt][j] = 0.f;
#pragma unroll
    for (int st = 0; st < 4; st++) {
      const f32x4 acc = mfma_tile(TW->w2t + 16 * vt * 64, 64, L.a1 + 16 * st * 64, 64, 64, zero);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int v = 16 * vt + 4 * g + j, s = 16 * st + n;
        const float dz = acc[j] * d1[vt][st][j];
        db1[vt][j] += dz;
        L.a2[v * 64 + s] = f2bf(dz);
      }
    }
  }
  wave_lds_sync();
  // dW1[u][k] = dZ1^T X
#pragma unroll
  for (int ut = 0; ut < 4; ut++) {
    const f32x4 acc = mfma_tile(L.a2 + 16 * ut * 64, 64, L.xt, 64, 64, zero);
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (n < kMlpIn) gp[kMlpW1 + (16 * ut + 4 * g + j) * kMlpIn + n] = acc[j];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const float v = sum16(db1[ut][j]);
      if (n == 0) gp[kMlpB1 + 16 * ut + 4 * g + j] = v;
    }
  }
  for (int o = 32; o > 0; o >>= 1) loss += __shfl_xor(loss, o);
  if (l == 0) lpart[blockIdx.x] = loss;
}

// sums the chunk gradients in chunk order; g[kMlpNParams] = loss
__global__ void __launch_bounds__(256) k_mlp_gsum(const float *gpart, int nparts,
                                                  const float *lpart, float *gsum) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < kMlpNParams) {
    float s = 0.f;
    for (int k = 0; k < nparts; k++) s += gpart[(long)k * kMlpNParams + i];
    gsum[i] = s;
  } else if (i == kMlpNParams) {
    float s = 0.f;
    for (int k = 0; k < nparts; k++) s += lpart[k];
    gsum[i] = s;
  }
}

// Adam on the FP32 master weights (gradient = chunk sum in chunk order),
// then the BF16 layouts; adds the step's loss times lscale to *eloss
__global__ void __launch_bounds__(256) k_mlp_adam(float *theta, float *am, float *av,
                                                  const float *gpart, int nparts, float lr,
                                                  float b1, float b2, float eps, float bc1,
                                                  float bc2, MlpWeights *W, MlpTrainWeights *T,
                                                  const float *lpart, double *eloss,
                                                  double lscale) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < kMlpNParams) {
    float gi = 0.f;
    for (int k = 0; k < nparts; k++) gi += gpart[(long)k * kMlpNParams + i];
    const float mi = b1 * am[i] + (1.f - b1) * gi;
    const float vi = b2 * av[i] + (1.f - b2) * gi * gi;
    am[i] = mi;
    av[i] = vi;
    const float th = theta[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    theta[i] = th;
    scatter_param(i, th, W, T);
  } else if (i == kMlpNParams && eloss) {
    double s = 0.0;
    for (int k = 0; k < nparts; k++) s += lpart[k];
    *eloss += s * lscale;
  }
}

// standardized FP32 inputs [n][16] -> BF16 queue rows (the layout the
// element pass writes)
__global__ void __launch_bounds__(256) k_mlp_pack_rows(const float *X, int n, uint16_t *rows) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float x[kMlpIn];
#pragma unroll
  for (int k = 0; k < kMlpIn; k++) x[k] = X[(long)i * 16 + k];
  mlp_store_inputs(rows + (size_t)i * kMlpRowHalves, x);
}

// The return map of the surrogate step on given points, in the step's
// three stages: the trial stress and the network inputs of every point into
// its queue row, the network on the rows, then the surrogate return (with
// its acceptance checks and FP64 fallback) where the trial stress yields.
__device__ __forceinline__ float rm_trial(const Params &p, int i, const float *de, const float *so,
                                          float d[6], float o[6], float s[6]) {
#pragma unroll
  for (int c = 0; c < 6; c++) {
    d[c] = de[(long)i * 6 + c];
    o[c] = so[(long)i * 6 + c];
  }
  return ep_trial<float>(p, d, o, s);
}

__global__ void __launch_bounds__(256) k_rm_prepare(Params p, int n, const float *de,
                                                    const float *so, const float *ep,
                                                    uint16_t *rows) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float d[6], o[6], s[6], x[kMlpIn];
  (void)rm_trial(p, i, de, so, d, o, s);
  mlp_inputs<float>(*p.mlp, d, o, ep[i], x);
  mlp_store_inputs(rows + (size_t)i * kMlpRowHalves, x);
}

__global__ void __launch_bounds__(256) k_rm_finish(Params p, int n, const float *de,
                                                   const float *so, const float *ep,
                                                   const uint16_t *rows, float *s_out,
                                                   float *ep_out, int *how) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  float d[6], o[6], s[6];
  (void)rm_trial(p, i, de, so, d, o, s);
  float e = ep[i];
  const float q = yield_overshoot<float>(p, s, e);
  int h = MLP_ELASTIC;
  if (q > 0.f) {
    float y[kMlpOut];
    mlp_load_outputs(rows + (size_t)i * kMlpRowHalves, y);
    mlp_apply<float>(p, y, q, e, s, h);
  }
#pragma unroll
  for (int c = 0; c < 6; c++) s_out[(long)i * 6 + c] = s[c];
  ep_out[i] = e;
  how[i] = h;
}

// ---------------------------------------------------------------------------
// per-context buffers
// ---------------------------------------------------------------------------

int mlp_alloc(MlpHost &m, int nstats, hipStream_t s, std::string &err) {
  m.nstats = nstats > 0 ? nstats : 1;
  const size_t sb = 2 * sizeof(unsigned long long) * (size_t)m.nstats;
  if (hipMalloc(&m.d_params, sizeof(MlpParams)) != hipSuccess ||
      hipMalloc(&m.d_weights, sizeof(MlpWeights)) != hipSuccess ||
      hipMalloc(&m.d_stats, sb) != hipSuccess) {
    err = "hipMalloc for the surrogate failed";
    return CPI_ERR_ALLOC;
  }
  if (hipMemsetAsync(m.d_stats, 0, sb, s) != hipSuccess ||
      hipMemsetAsync(m.d_weights, 0, sizeof(MlpWeights), s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    err = "hipMemset for the surrogate failed";
    return CPI_ERR_HIP;
  }
  return CPI_OK;
}

int mlp_samples_alloc(MlpHost &m, int cap, hipStream_t s, std::string &err) {
  if (m.d_samples) (void)hipFree(m.d_samples);
  m.d_samples = nullptr;
  m.sample_cap = 0;
  if (!m.d_sample_count && hipMalloc(&m.d_sample_count, sizeof(int)) != hipSuccess) {
    err = "hipMalloc for the sample counter failed";
    return CPI_ERR_ALLOC;
  }
  if (hipMemsetAsync(m.d_sample_count, 0, sizeof(int), s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    err = "hipMemset of the sample counter failed";
    return CPI_ERR_HIP;
  }
  if (cap > 0) {
    if (hipMalloc(&m.d_samples, sizeof(float) * kSampleFloats * (size_t)cap) != hipSuccess) {
      err = "hipMalloc of " + std::to_string(cap) + " samples failed";
      return CPI_ERR_ALLOC;
    }
    m.sample_cap = cap;
  }
  return CPI_OK;
}

void mlp_free(MlpHost &m) {
  if (m.d_params) (void)hipFree(m.d_params);
  if (m.d_stats) (void)hipFree(m.d_stats);
  if (m.d_weights) (void)hipFree(m.d_weights);
  if (m.d_samples) (void)hipFree(m.d_samples);
  if (m.d_sample_count) (void)hipFree(m.d_sample_count);
  m = MlpHost{};
}

MlpMaterial mlp_material(const cpi_material &mat) {
  const double k = mat.E / ((1.0 + mat.nu) * (1.0 - 2.0 * mat.nu));
  MlpMaterial r;
  r.c11 = (1.0 - mat.nu) * k;
  r.c12 = mat.nu * k;
  r.c44 = 0.5 * (1.0 - 2.0 * mat.nu) * k;
  r.mu = mat.E / (2.0 * (1.0 + mat.nu));
  r.sy0 = mat.Y0;
  r.H = mat.H;
  return r;
}

}  // namespace cpi

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

using namespace cpi;

namespace {

#define MCHK(m, expr)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess) {                                                             \
      (m)->err = std::string(#expr) + " failed: " + hipGetErrorString(e_);              \
      return CPI_ERR_HIP;                                                               \
    }                                                                                   \
  } while (0)

int mfail(cpi_mlp *m, int code, const std::string &msg) {
  m->err = msg;
  return code;
}

template <typename T>
int mdalloc(cpi_mlp *m, T **p, size_t n) {
  void *q = nullptr;
  if (hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess)
    return mfail(m, CPI_ERR_ALLOC, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed");
  *p = (T *)q;
  return CPI_OK;
}

uint64_t splitmix(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double u01(uint64_t &s) { return (splitmix(s) >> 11) * (1.0 / 9007199254740992.0); }

// minimum trial overshoot (relative to sy0) of a training sample: below it
// the scaled targets are dominated by the sample's own rounding
constexpr double kMinOvershoot = 1e-3;

// raw inputs and targets of one sample (de, sigma_n, ep_n, sigma_{n+1},
// ep_{n+1}); false if its trial overshoot is below kMinOvershoot * sy0
bool sample_xt(const MlpMaterial &mm, const float *smp, double x[kMlpIn], double t[kMlpOut]) {
  double de[6], so[6], tr[6];
  for (int c = 0; c < 6; c++) {
    de[c] = smp[c];
    so[c] = smp[6 + c];
  }
  tr[0] = so[0] + mm.c11 * de[0] + mm.c12 * (de[1] + de[2]);
  tr[1] = so[1] + mm.c11 * de[1] + mm.c12 * (de[0] + de[2]);
  tr[2] = so[2] + mm.c11 * de[2] + mm.c12 * (de[0] + de[1]);
  tr[3] = so[3] + mm.c44 * de[3];
  tr[4] = so[4] + mm.c44 * de[4];
  tr[5] = so[5] + mm.c44 * de[5];
  const double p = (tr[0] + tr[1] + tr[2]) / 3.0;
  const double d0 = tr[0] - p, d1 = tr[1] - p, d2 = tr[2] - p;
  const double nrm = std::sqrt(d0 * d0 + d1 * d1 + d2 * d2 +
                               2.0 * (tr[3] * tr[3] + tr[4] * tr[4] + tr[5] * tr[5]));
  const double ep0 = smp[12], ep1 = smp[19];
  const double q = nrm - std::sqrt(2.0 / 3.0) * (mm.sy0 + mm.H * ep0);
  if (!(q > kMinOvershoot * mm.sy0) || !std::isfinite(q)) return false;
  for (int c = 0; c < 6; c++) {
    x[c] = de[c];
    x[6 + c] = so[c];
    t[c] = (tr[c] - (double)smp[13 + c]) / q;
  }
  x[12] = ep0;
  t[6] = (ep1 - ep0) * mm.sy0 / q;
  return true;
}

void stat_of(const std::vector<double> &v, int dim, int n, int c, float &mean, float &sd) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += v[(size_t)i * dim + c];
  const double mu = n ? s / n : 0.0;
  double q = 0.0;
  for (int i = 0; i < n; i++) {
    const double d = v[(size_t)i * dim + c] - mu;
    q += d * d;
  }
  double sdv = n ? std::sqrt(q / n) : 1.0;
  const double floor = 1e-3 * std::fabs(mu);
  if (sdv < floor) sdv = floor;
  if (!(sdv > 0.0)) sdv = 1.0;
  mean = (float)mu;
  sd = (float)sdv;
}

int upload_view(cpi_mlp *m) {
  MlpParams h{};
  h.weights = m->W;
  h.w1 = m->W->w1;
  h.w2p = m->W->w2p;
  h.w3p = m->W->w3p;
  h.b1 = m->W->b1;
  h.b2 = m->W->b2;
  h.b3 = m->W->b3;
  for (int k = 0; k < kMlpIn; k++) {
    h.in_mean[k] = m->in_mean[k];
    h.in_istd[k] = 1.0f / m->in_std[k];
  }
  for (int k = 0; k < kMlpOut; k++) {
    h.out_mean[k] = m->out_mean[k];
    h.out_std[k] = m->out_std[k];
  }
  h.c11 = (float)m->mm.c11;
  h.c12 = (float)m->mm.c12;
  h.c44 = (float)m->mm.c44;
  h.sy0 = (float)m->mm.sy0;
  h.H = (float)m->mm.H;
  MCHK(m, hipMemcpyAsync(m->dP, &h, sizeof(h), hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return CPI_OK;
}

// theta on the device -> BF16 layouts; Adam state cleared
int install_theta(cpi_mlp *m, const std::vector<float> &theta) {
  MCHK(m, hipMemcpyAsync(m->theta, theta.data(), sizeof(float) * kMlpNParams,
                         hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipMemsetAsync(m->adam_m, 0, sizeof(float) * kMlpNParams, m->stream));
  MCHK(m, hipMemsetAsync(m->adam_v, 0, sizeof(float) * kMlpNParams, m->stream));
  hipLaunchKernelGGL(k_mlp_scatter, dim3((kMlpNParams + 255) / 256), dim3(256), 0, m->stream,
                     m->theta, m->W, m->TW);
  MCHK(m, hipGetLastError());
  MCHK(m, hipStreamSynchronize(m->stream));
  m->adam_t = 0;
  return CPI_OK;
}

int ensure_data(cpi_mlp *m, size_t n) {
  if (m->data_cap >= n) return CPI_OK;
  if (m->dX) (void)hipFree(m->dX);
  if (m->dT) (void)hipFree(m->dT);
  if (m->dIdx) (void)hipFree(m->dIdx);
  m->dX = m->dT = nullptr;
  m->dIdx = nullptr;
  m->data_cap = 0;
  int rc;
  // dT doubles as the inference queue, whose last tile may read 15 rows past n
  if ((rc = mdalloc(m, &m->dX, 16 * n)) || (rc = mdalloc(m, &m->dT, 8 * (n + 16))) ||
      (rc = mdalloc(m, &m->dIdx, n)))
    return rc;
  m->data_cap = n;
  return CPI_OK;
}

// host [n][13] / [n][7] -> padded device rows
int upload_rows(cpi_mlp *m, const float *x, const float *t, int n) {
  int rc = ensure_data(m, (size_t)n);
  if (rc) return rc;
  std::vector<float> hx((size_t)16 * n, 0.f), ht((size_t)8 * n, 0.f);
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < kMlpIn; k++) hx[(size_t)i * 16 + k] = x[(size_t)i * kMlpIn + k];
    if (t)
      for (int k = 0; k < kMlpOut; k++) ht[(size_t)i * 8 + k] = t[(size_t)i * kMlpOut + k];
  }
  MCHK(m, hipMemcpyAsync(m->dX, hx.data(), sizeof(float) * hx.size(), hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipMemcpyAsync(m->dT, ht.data(), sizeof(float) * ht.size(), hipMemcpyHostToDevice, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return CPI_OK;
}

// samples -> standardized (x, t) of the kept ones
int standardize(cpi_mlp *m, const float *samples, int n, std::vector<float> &x,
                std::vector<float> &t, int &kept, bool fit) {
  std::vector<double> rx, rt;
  rx.reserve((size_t)n * kMlpIn);
  rt.reserve((size_t)n * kMlpOut);
  kept = 0;
  for (int i = 0; i < n; i++) {
    double xi[kMlpIn], ti[kMlpOut];
    if (!sample_xt(m->mm, samples + (size_t)i * kSampleFloats, xi, ti)) continue;
    rx.insert(rx.end(), xi, xi + kMlpIn);
    rt.insert(rt.end(), ti, ti + kMlpOut);
    kept++;
  }
  if (fit) {
    if (kept < 1) return mfail(m, CPI_ERR_INVALID, "no sample with a trial overshoot above 1e-3 sy0");
    for (int c = 0; c < kMlpIn; c++) stat_of(rx, kMlpIn, kept, c, m->in_mean[c], m->in_std[c]);
    for (int c = 0; c < kMlpOut; c++) stat_of(rt, kMlpOut, kept, c, m->out_mean[c], m->out_std[c]);
    m->scaled = true;
    int rc = upload_view(m);
    if (rc) return rc;
  }
  x.resize((size_t)kept * kMlpIn);
  t.resize((size_t)kept * kMlpOut);
  for (int i = 0; i < kept; i++) {
    for (int c = 0; c < kMlpIn; c++)
      x[(size_t)i * kMlpIn + c] = (float)((rx[(size_t)i * kMlpIn + c] - m->in_mean[c]) / m->in_std[c]);
    for (int c = 0; c < kMlpOut; c++)
      t[(size_t)i * kMlpOut + c] =
          (float)((rt[(size_t)i * kMlpOut + c] - m->out_mean[c]) / m->out_std[c]);
  }
  return CPI_OK;
}

}  // namespace

extern "C" void cpi_mlp_default_train_opts(cpi_mlp_train_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->epochs = 60;
  o->batch = 1024;
  o->lr = 2e-3f;
  o->lr_final = 2e-5f;
  o->beta1 = 0.9f;
  o->beta2 = 0.999f;
  o->eps = 1e-8f;
  o->seed = 1;
}

extern "C" int cpi_mlp_create(int32_t device, uint64_t seed, const cpi_material *mat,
                              cpi_mlp **out) {
  if (!out || !mat) {
    set_global_error("cpi_mlp_create: mat and out are required");
    return CPI_ERR_INVALID;
  }
  *out = nullptr;
  if (!(mat->E > 0.0) || !(mat->nu >= 0.0 && mat->nu < 0.5) || !(mat->Y0 > 0.0) ||
      !(mat->H >= 0.0)) {
    set_global_error("cpi_mlp_create: needs E > 0, 0 <= nu < 0.5, Y0 > 0, H >= 0");
    return CPI_ERR_INVALID;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    set_global_error("cpi_mlp_create: no such HIP device");
    return CPI_ERR_NO_DEVICE;
  }
  hipDeviceProp_t prop;
  if (hipSetDevice(device) != hipSuccess || hipGetDeviceProperties(&prop, device) != hipSuccess) {
    set_global_error("cpi_mlp_create: cannot query the device");
    return CPI_ERR_HIP;
  }
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_global_error(std::string("cpi_mlp_create: device is ") + prop.gcnArchName +
                     ", this library is built for gfx950 only");
    return CPI_ERR_NO_DEVICE;
  }
  cpi_mlp *m = new cpi_mlp;
  m->dev = device;
  m->mat = *mat;
  m->mm = mlp_material(*mat);
  for (int k = 0; k < kMlpIn; k++) {
    m->in_mean[k] = 0.f;
    m->in_std[k] = 1.f;
  }
  for (int k = 0; k < kMlpOut; k++) {
    m->out_mean[k] = 0.f;
    m->out_std[k] = 1.f;
  }
  int rc = CPI_OK;
  auto bail = [&](int code) {
    set_global_error("cpi_mlp_create: " + m->err);
    cpi_mlp_destroy(m);
    return code;
  };
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(mfail(m, CPI_ERR_HIP, "stream creation failed"));
  if ((rc = mdalloc(m, &m->theta, kMlpNParams)) || (rc = mdalloc(m, &m->adam_m, kMlpNParams)) ||
      (rc = mdalloc(m, &m->adam_v, kMlpNParams)) || (rc = mdalloc(m, &m->W, 1)) ||
      (rc = mdalloc(m, &m->TW, 1)) || (rc = mdalloc(m, &m->dP, 1)) ||
      (rc = mdalloc(m, &m->gpart, (size_t)kMaxParts * kMlpNParams)) ||
      (rc = mdalloc(m, &m->lpart, kMaxParts)) || (rc = mdalloc(m, &m->gsum, kMlpNParams + 1)))
    return bail(rc);
  m->part_cap = kMaxParts;
  if (hipMemsetAsync(m->W, 0, sizeof(MlpWeights), m->stream) != hipSuccess ||
      hipMemsetAsync(m->TW, 0, sizeof(MlpTrainWeights), m->stream) != hipSuccess ||
      hipStreamSynchronize(m->stream) != hipSuccess)
    return bail(mfail(m, CPI_ERR_HIP, "hipMemset failed"));
  // Glorot-uniform weights, zero biases
  std::vector<float> theta(kMlpNParams, 0.f);
  uint64_t s = seed;
  auto fill = [&](int off, int rows, int cols) {
    const double a = std::sqrt(6.0 / (rows + cols));
    for (int i = 0; i < rows * cols; i++) theta[off + i] = (float)((2.0 * u01(s) - 1.0) * a);
  };
  fill(kMlpW1, kMlpHid, kMlpIn);
  fill(kMlpW2, kMlpHid, kMlpHid);
  fill(kMlpW3, kMlpOut, kMlpHid);
  if ((rc = install_theta(m, theta)) || (rc = upload_view(m))) return bail(rc);
  *out = m;
  return CPI_OK;
}

extern "C" void cpi_mlp_destroy(cpi_mlp *m) {
  if (!m) return;
  (void)hipSetDevice(m->dev);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  void *ptrs[] = {m->theta, m->adam_m, m->adam_v, m->W,  m->TW,   m->dP,   m->gpart,
                  m->lpart, m->gsum,   m->dX,     m->dT, m->dIdx, m->dLoss};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
}

extern "C" const char *cpi_mlp_last_error(const cpi_mlp *m) {
  return m ? m->err.c_str() : global_error();
}

extern "C" int cpi_mlp_set_params(cpi_mlp *m, const float *params) {
  if (!m || !params) return CPI_ERR_INVALID;
  for (int i = 0; i < kMlpNParams; i++)
    if (!std::isfinite(params[i])) return mfail(m, CPI_ERR_INVALID, "non-finite parameter");
  (void)hipSetDevice(m->dev);
  return install_theta(m, std::vector<float>(params, params + kMlpNParams));
}

extern "C" int cpi_mlp_get_params(cpi_mlp *m, float *params) {
  if (!m || !params) return CPI_ERR_INVALID;
  (void)hipSetDevice(m->dev);
  MCHK(m, hipMemcpyAsync(params, m->theta, sizeof(float) * kMlpNParams, hipMemcpyDeviceToHost,
                         m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  return CPI_OK;
}

extern "C" int cpi_mlp_set_scaling(cpi_mlp *m, const float *in_mean, const float *in_std,
                                   const float *out_mean, const float *out_std) {
  if (!m || !in_mean || !in_std || !out_mean || !out_std) return CPI_ERR_INVALID;
  for (int k = 0; k < kMlpIn; k++)
    if (!(in_std[k] > 0.f) || !std::isfinite(in_mean[k]) || !std::isfinite(in_std[k]))
      return mfail(m, CPI_ERR_INVALID, "input scales must be finite and positive");
  for (int k = 0; k < kMlpOut; k++)
    if (!(out_std[k] > 0.f) || !std::isfinite(out_mean[k]) || !std::isfinite(out_std[k]))
      return mfail(m, CPI_ERR_INVALID, "output scales must be finite and positive");
  std::memcpy(m->in_mean, in_mean, sizeof(m->in_mean));
  std::memcpy(m->in_std, in_std, sizeof(m->in_std));
  std::memcpy(m->out_mean, out_mean, sizeof(m->out_mean));
  std::memcpy(m->out_std, out_std, sizeof(m->out_std));
  m->scaled = true;
  (void)hipSetDevice(m->dev);
  return upload_view(m);
}

extern "C" int cpi_mlp_get_scaling(cpi_mlp *m, float *in_mean, float *in_std, float *out_mean,
                                   float *out_std) {
  if (!m) return CPI_ERR_INVALID;
  if (in_mean) std::memcpy(in_mean, m->in_mean, sizeof(m->in_mean));
  if (in_std) std::memcpy(in_std, m->in_std, sizeof(m->in_std));
  if (out_mean) std::memcpy(out_mean, m->out_mean, sizeof(m->out_mean));
  if (out_std) std::memcpy(out_std, m->out_std, sizeof(m->out_std));
  return CPI_OK;
}

extern "C" int cpi_mlp_fit_scaling(cpi_mlp *m, const float *samples, int32_t n) {
  if (!m || !samples || n < 1) return CPI_ERR_INVALID;
  (void)hipSetDevice(m->dev);
  std::vector<float> x, t;
  int kept = 0;
  return standardize(m, samples, n, x, t, kept, true);
}

extern "C" int cpi_mlp_targets(cpi_mlp *m, const float *samples, int32_t n, float *x, float *t,
                               int32_t *kept) {
  if (!m || !samples || n < 0 || !kept) return CPI_ERR_INVALID;
  std::vector<float> hx, ht;
  int k = 0;
  int rc = standardize(m, samples, n, hx, ht, k, false);
  if (rc) return rc;
  if (x) std::memcpy(x, hx.data(), sizeof(float) * hx.size());
  if (t) std::memcpy(t, ht.data(), sizeof(float) * ht.size());
  *kept = k;
  return CPI_OK;
}

extern "C" int cpi_mlp_forward(cpi_mlp *m, const float *x, int32_t n, float *y) {
  if (!m || !x || !y || n < 0) return CPI_ERR_INVALID;
  if (n == 0) return CPI_OK;
  (void)hipSetDevice(m->dev);
  int rc = upload_rows(m, x, nullptr, n);
  if (rc) return rc;
  // dT's [rows][8] floats are the queue rows: BF16 inputs in, outputs back
  uint16_t *rows = reinterpret_cast<uint16_t *>(m->dT);
  hipLaunchKernelGGL(k_mlp_pack_rows, dim3((n + 255) / 256), dim3(256), 0, m->stream, m->dX, n,
                     rows);
  MCHK(m, hipGetLastError());
  MCHK(m, launch_mlp_rows(m->dP, rows, n, m->stream));
  std::vector<float> hy((size_t)8 * n);
  MCHK(m, hipMemcpyAsync(hy.data(), m->dT, sizeof(float) * hy.size(), hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  for (int i = 0; i < n; i++)
    for (int k = 0; k < kMlpOut; k++) y[(size_t)i * kMlpOut + k] = hy[(size_t)i * 8 + k];
  return CPI_OK;
}

extern "C" int cpi_mlp_gradient(cpi_mlp *m, const float *x, const float *t, int32_t n,
                                float *grad, float *loss) {
  if (!m || !x || !t || n < 1) return CPI_ERR_INVALID;
  const int parts = (n + 63) / 64;
  if (parts > m->part_cap)
    return mfail(m, CPI_ERR_INVALID, "cpi_mlp_gradient: at most " + std::to_string(64 * m->part_cap) + " samples");
  (void)hipSetDevice(m->dev);
  int rc = upload_rows(m, x, t, n);
  if (rc) return rc;
  hipLaunchKernelGGL(k_mlp_grad, dim3(parts), dim3(64), 0, m->stream, m->dX, m->dT,
                     (const int *)nullptr, 0, n, 1.0f / n, m->W, m->TW, m->gpart, m->lpart);
  MCHK(m, hipGetLastError());
  hipLaunchKernelGGL(k_mlp_gsum, dim3((kMlpNParams + 1 + 255) / 256), dim3(256), 0, m->stream,
                     m->gpart, parts, m->lpart, m->gsum);
  MCHK(m, hipGetLastError());
  std::vector<float> h(kMlpNParams + 1);
  MCHK(m, hipMemcpyAsync(h.data(), m->gsum, sizeof(float) * h.size(), hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  if (grad) std::memcpy(grad, h.data(), sizeof(float) * kMlpNParams);
  if (loss) *loss = h[kMlpNParams];
  return CPI_OK;
}

extern "C" int cpi_mlp_train(cpi_mlp *m, const float *samples, int32_t n,
                             const cpi_mlp_train_opts *opt, float *loss_history) {
  if (!m || !samples || n < 1) return CPI_ERR_INVALID;
  cpi_mlp_train_opts o;
  if (opt)
    o = *opt;
  else
    cpi_mlp_default_train_opts(&o);
  if (o.epochs < 1 || o.batch < 1 || o.batch > 64 * m->part_cap || !(o.lr > 0.f) ||
      !(o.lr_final > 0.f) || !(o.beta1 >= 0.f && o.beta1 < 1.f) ||
      !(o.beta2 >= 0.f && o.beta2 < 1.f) || !(o.eps > 0.f))
    return mfail(m, CPI_ERR_INVALID, "cpi_mlp_train: invalid options (batch at most " +
                                         std::to_string(64 * m->part_cap) + ")");
  (void)hipSetDevice(m->dev);
  std::vector<float> x, t;
  int kept = 0;
  int rc = standardize(m, samples, n, x, t, kept, !m->scaled);
  if (rc) return rc;
  if (kept < 1) return mfail(m, CPI_ERR_INVALID, "no sample with a trial overshoot above 1e-3 sy0");
  if ((rc = upload_rows(m, x.data(), t.data(), kept))) return rc;
  if (m->loss_cap < o.epochs) {
    if (m->dLoss) (void)hipFree(m->dLoss);
    m->dLoss = nullptr;
    m->loss_cap = 0;
    if ((rc = mdalloc(m, &m->dLoss, o.epochs))) return rc;
    m->loss_cap = o.epochs;
  }
  MCHK(m, hipMemsetAsync(m->dLoss, 0, sizeof(double) * o.epochs, m->stream));
  std::vector<int> perm(kept);
  for (int i = 0; i < kept; i++) perm[i] = i;
  uint64_t rs = o.seed;
  const double pi = 3.14159265358979323846;
  for (int ep = 0; ep < o.epochs; ep++) {
    // the previous epoch's upload of perm must be done before it changes
    MCHK(m, hipStreamSynchronize(m->stream));
    for (int i = kept - 1; i > 0; i--) std::swap(perm[i], perm[splitmix(rs) % (uint64_t)(i + 1)]);
    MCHK(m, hipMemcpyAsync(m->dIdx, perm.data(), sizeof(int) * kept, hipMemcpyHostToDevice, m->stream));
    const double cosf_ = 0.5 * (1.0 + std::cos(pi * ep / o.epochs));
    const float lr = (float)(o.lr_final + (o.lr - o.lr_final) * cosf_);
    for (int b0 = 0; b0 < kept; b0 += o.batch) {
      const int bn = std::min(o.batch, kept - b0), parts = (bn + 63) / 64;
      hipLaunchKernelGGL(k_mlp_grad, dim3(parts), dim3(64), 0, m->stream, m->dX, m->dT,
                         (const int *)m->dIdx, b0, bn, 1.0f / bn, m->W, m->TW, m->gpart, m->lpart);
      MCHK(m, hipGetLastError());
      m->adam_t++;
      const float bc1 = (float)(1.0 - std::pow((double)o.beta1, (double)m->adam_t));
      const float bc2 = (float)(1.0 - std::pow((double)o.beta2, (double)m->adam_t));
      hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpNParams + 1 + 255) / 256), dim3(256), 0, m->stream,
                         m->theta, m->adam_m, m->adam_v, (const float *)m->gpart, parts, lr,
                         o.beta1, o.beta2, o.eps, bc1, bc2, m->W, m->TW, (const float *)m->lpart,
                         m->dLoss + ep, (double)bn / kept);
      MCHK(m, hipGetLastError());
    }
  }
  std::vector<double> hl(o.epochs);
  MCHK(m, hipMemcpyAsync(hl.data(), m->dLoss, sizeof(double) * o.epochs, hipMemcpyDeviceToHost, m->stream));
  MCHK(m, hipStreamSynchronize(m->stream));
  if (loss_history)
    for (int i = 0; i < o.epochs; i++) loss_history[i] = (float)hl[i];
  return CPI_OK;
}

extern "C" int cpi_mlp_return_map(cpi_mlp *m, double tol, int32_t n, const float *de,
                                  const float *so, const float *ep, float *s_out, float *ep_out,
                                  int32_t *how) {
  if (!m || n < 0 || !de || !so || !ep || !s_out || !ep_out || !how || !(tol >= 0.0))
    return CPI_ERR_INVALID;
  if (n == 0) return CPI_OK;
  (void)hipSetDevice(m->dev);
  float *d = nullptr;
  int *dh = nullptr;
  // de, so, ep, s_out, ep_out, then the queue rows (8 floats each, one
  // tile of padding)
  const size_t nf = (size_t)n * 20 + ((size_t)n + 16) * 8;
  int rc;
  if ((rc = mdalloc(m, &d, nf))) return rc;
  if ((rc = mdalloc(m, &dh, (size_t)n))) {
    (void)hipFree(d);
    return rc;
  }
  float *dde = d, *dso = d + 6L * n, *dep = d + 12L * n, *ds = d + 13L * n, *dept = d + 19L * n;
  uint16_t *rows = reinterpret_cast<uint16_t *>(d + 20L * n);
  Params p{};
  p.c11 = m->mm.c11;
  p.c12 = m->mm.c12;
  p.c44 = m->mm.c44;
  p.mu = m->mm.mu;
  p.sy0 = m->mm.sy0;
  p.H = m->mm.H;
  p.mlp = m->dP;
  p.mlp_tol = tol;
  const int grid = (n + 255) / 256;
  hipError_t e = hipMemsetAsync(rows, 0, ((size_t)n + 16) * 32, m->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dde, de, sizeof(float) * 6 * n, hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dso, so, sizeof(float) * 6 * n, hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(dep, ep, sizeof(float) * n, hipMemcpyHostToDevice, m->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_rm_prepare, dim3(grid), dim3(256), 0, m->stream, p, n,
                       (const float *)dde, (const float *)dso, (const float *)dep, rows);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = launch_mlp_rows(m->dP, rows, n, m->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_rm_finish, dim3(grid), dim3(256), 0, m->stream, p, n, (const float *)dde,
                       (const float *)dso, (const float *)dep, (const uint16_t *)rows, ds, dept,
                       dh);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(s_out, ds, sizeof(float) * 6 * n, hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(ep_out, dept, sizeof(float) * n, hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(how, dh, sizeof(int) * n, hipMemcpyDeviceToHost, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  (void)hipFree(d);
  (void)hipFree(dh);
  if (e != hipSuccess) return mfail(m, CPI_ERR_HIP, std::string("cpi_mlp_return_map: ") + hipGetErrorString(e));
  return CPI_OK;
}
