// Host side of the MLP surrogate: device buffers for the network weights,
// the sample buffer and the fallback counters.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "mlp.h"

namespace cpi {

struct MlpHost {
  MlpParams *d_params = nullptr;  // device copy of the inference view
  int *d_stats = nullptr;         // [0] accepted, [1] fell back
  void *d_weights = nullptr;      // bf16 matrices + f32 biases
  bool ready = false;             // weights installed
};

int mlp_alloc(MlpHost &m, std::string &err);
void mlp_free(MlpHost &m);

}  // namespace cpi
