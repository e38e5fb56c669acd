// MLP surrogate of the J2 return map: device buffers, weight upload.
#include <hip/hip_runtime.h>

#include "../../include/cppimpact_hip.h"
#include "mlp_host.h"

namespace cpi {

int mlp_alloc(MlpHost &m, std::string &err) {
  if (hipMalloc(&m.d_params, sizeof(MlpParams)) != hipSuccess ||
      hipMalloc(&m.d_stats, 2 * sizeof(int)) != hipSuccess) {
    err = "hipMalloc for the surrogate failed";
    return CPI_ERR_ALLOC;
  }
  (void)hipMemset(m.d_stats, 0, 2 * sizeof(int));
  return CPI_OK;
}

void mlp_free(MlpHost &m) {
  if (m.d_params) (void)hipFree(m.d_params);
  if (m.d_stats) (void)hipFree(m.d_stats);
  if (m.d_weights) (void)hipFree(m.d_weights);
  m = MlpHost{};
}

}  // namespace cpi
