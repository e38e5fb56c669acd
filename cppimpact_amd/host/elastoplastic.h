// Elastoplastic<T, dof_per_node> (elastoplastic.h:6-23): same constructor and
// protected members; the read-only accessors are the one addition the HIP
// backend needs (the reference's Dynamics never reads beta, H or Y0).
#pragma once

#include <stdexcept>

#include "basematerial.h"

template <typename T, int dof_per_node>
class Elastoplastic : public BaseMaterial<T, dof_per_node> {
 public:
  Elastoplastic(T E, T rho, T nu, T beta, T H, T Y0, std::string name = "Material")
      : BaseMaterial<T, dof_per_node>("InfinitesimalElastoplastic", E, rho, nu, name),
        beta(beta),
        H(H),
        Y0_initial_yield_stress(Y0) {}

  virtual ~Elastoplastic() {}

  T get_beta() const { return beta; }
  T get_H() const { return H; }
  T get_Y0() const { return Y0_initial_yield_stress; }

 protected:
  T beta;
  T H;
  T Y0_initial_yield_stress;
};
