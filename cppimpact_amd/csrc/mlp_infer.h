// Surrogate return map used by the PHYS_EP_MLP element kernels.
#pragma once

#include "device.h"
#include "mlp.h"

namespace cpi {

template <typename R>
__device__ R mlp_return(const Params &p, const R de[6], const R so[6], R &ep, R s[6]) {
  (void)de;
  (void)so;
  return radial_return<R>(p, s, ep);
}

}  // namespace cpi
