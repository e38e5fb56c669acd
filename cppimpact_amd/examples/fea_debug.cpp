// The reference's known-answer driver (FEA_debug.cu:30-130) on the HIP
// backend: load a mesh, apply the prescribed displacement field and print
// the total strain energy and volume computed on the GPU.
//   ./fea_debug_<basis> MESH.inp [deformation 0|1]
// (the reference picks the mesh by an index into its input/ directory; here
// the path is given directly).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../host/analysis.h"
#include "../host/cppimpact_defs.h"
#include "../host/elastoplastic.h"
#include "../host/mesh.h"
#include "../host/physics.h"
#include "../host/tetrahedral.h"
#include "../host/wall.h"
#include "../host/dynamics_hip.h"

using T = double;
#if defined(USE_LINEAR_BASIS)
using Basis = TetrahedralBasisLinear<T>;
#elif defined(USE_QUADRATIC_BASIS)
using Basis = TetrahedralBasisQuadratic<T>;
#else
#error "No basis type defined"
#endif

int main(int argc, char *argv[]) {
  using Quadrature = TetrahedralQuadrature5pts;
  using Physics = NeohookeanPhysics<T>;
  using Analysis = FEAnalysis<T, Basis, Quadrature, Physics>;
  constexpr int dof_per_node = 3;
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s MESH.inp [deformation 0|1] [vtk dir]\n", argv[0]);
    return 1;
  }
  std::string filename(argv[1]);
  const int deformation = argc >= 3 ? std::atoi(argv[2]) : 0;

  Mesh<T, Basis::nodes_per_element> tensile;
  T E = 68.9E9;  // Pa
  T rho = 2700;  // kg/m3
  T nu = 0.33;
  T beta = 0.0;
  T H = 10;
  T Y0 = 1.9 * std::sqrt(3.0);
  std::string name = "AL6061";

  Elastoplastic<T, dof_per_node> material(E, rho, nu, beta, H, Y0, name);
  if (tensile.load_mesh(filename) != 0) return 1;

  const int normal = 1;
  std::string wall_name = "Wall";
  T location = -10;
  Wall<T, 2, Basis> w(wall_name, location, E * 10, tensile.slave_nodes, tensile.num_slave_nodes,
                      normal);

  Dynamics<T, Basis, Analysis, Quadrature> dyna(&tensile, &material, &w);
  if (argc >= 4) dyna.set_export_directory(argv[3]);

  auto start = std::chrono::high_resolution_clock::now();
  dyna.debug_strain(0.01, deformation);
  auto end = std::chrono::high_resolution_clock::now();
  std::chrono::duration<double> elapsed = end - start;
  std::cout << "Elapsed time: " << elapsed.count() << " seconds" << std::endl;
  return 0;
}
