// Host side of the MLP surrogate: the trainable network object (cpi_mlp)
// and the per-context buffers the element kernels read (MlpHost).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/cppimpact_hip.h"
#include "mlp.h"

namespace cpi {

// Per-context surrogate state: a copy of the installed weights, the
// inference view, the accept/fallback counters and the sample buffer.
struct MlpHost {
  MlpParams *d_params = nullptr;       // device copy of the inference view
  MlpWeights *d_weights = nullptr;     // installed weights
  unsigned long long *d_stats = nullptr;  // [nstats][2] accepted, fell back
  int nstats = 0;                      // one pair per element block
  bool ready = false;                  // weights installed
  float *d_samples = nullptr;          // [cap][kSampleFloats]
  int *d_sample_count = nullptr;
  int sample_cap = 0;
};

void set_global_error(const std::string &s);
const char *global_error();

int mlp_alloc(MlpHost &m, int nstats, hipStream_t s, std::string &err);
void mlp_free(MlpHost &m);
int mlp_samples_alloc(MlpHost &m, int cap, hipStream_t s, std::string &err);

// material constants the network is tied to (D and the yield curve)
struct MlpMaterial {
  double c11, c12, c44, mu, sy0, H;
};
MlpMaterial mlp_material(const cpi_material &mat);

}  // namespace cpi

struct cpi_mlp {
  int dev = 0;
  hipStream_t stream = nullptr;
  std::string err;
  cpi_material mat{};
  cpi::MlpMaterial mm{};
  float in_mean[cpi::kMlpIn], in_std[cpi::kMlpIn], out_mean[cpi::kMlpOut], out_std[cpi::kMlpOut];
  bool scaled = false;
  float *theta = nullptr, *adam_m = nullptr, *adam_v = nullptr;  // FP32 master + Adam moments
  int64_t adam_t = 0;
  cpi::MlpWeights *W = nullptr;
  cpi::MlpTrainWeights *TW = nullptr;
  cpi::MlpParams *dP = nullptr;  // inference view of W (forward / return-map tests)
  float *gpart = nullptr, *lpart = nullptr, *gsum = nullptr;
  int part_cap = 0;
  float *dX = nullptr, *dT = nullptr;  // standardized inputs [n][16] and targets [n][8]
  int *dIdx = nullptr;
  size_t data_cap = 0;
  double *dLoss = nullptr;  // per-epoch loss accumulator
  int loss_cap = 0;
};
