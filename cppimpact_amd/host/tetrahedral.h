// Basis, quadrature, physics and analysis types of the reference as type
// tags: the HIP backend evaluates them on the GPU, the host side only needs
// their sizes and which rule they name (tetrahedral.h:7-74, 182-269,
// 521-567; physics.h; analysis.h).
#pragma once

#include "cppimpact_defs.h"

template <typename T>
class TetrahedralBasisLinear {
 public:
  static constexpr int spatial_dim = 3;
  static constexpr int nodes_per_element = 4;
};

template <typename T>
class TetrahedralBasisQuadratic {
 public:
  static constexpr int spatial_dim = 3;
  static constexpr int nodes_per_element = 10;
};

// the reference's 5-point rule with its truncated literals
class TetrahedralQuadrature5pts {
 public:
  static constexpr int num_quadrature_pts = 5;
  static constexpr int cpi_rule = CPI_QUAD_REF5;
};

// the reference's 4-point rule with its truncated literals
class TetrahedralQuadrature4pts {
 public:
  static constexpr int num_quadrature_pts = 4;
  static constexpr int cpi_rule = CPI_QUAD_REF4;
};

template <typename T>
class NeohookeanPhysics {};

template <typename T, class Basis, class Quadrature, class Physics>
class FEAnalysis {
 public:
  static constexpr int spatial_dim = Basis::spatial_dim;
  static constexpr int nodes_per_element = Basis::nodes_per_element;
};
