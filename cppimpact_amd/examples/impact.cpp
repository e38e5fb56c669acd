// The reference's dynamic driver (test.cu:30-106) on the HIP backend: the
// same classes, constructor calls and solve(); only the include that picks
// the backend differs.  Extra flags (all optional) make it usable as a test
// and benchmark driver:
//   --smoke            10 steps (test.cu:44-48)
//   --mesh FILE        .inp mesh (default: the reference's fuselage mesh)
//   --velocity VZ      initial z velocity (default -10)
//   --wall Z           wall location (default 0.15 - 0.00005)
//   --dt DT --tend T   time step and end time (defaults 1e-7, 0.5)
//   --ep               J2 elastoplastic force model instead of the reference's
//   --dump PREFIX      write PREFIX_x.txt / PREFIX_v.txt (final state, %.17g)
//   --out DIR          VTK export directory (default ../gpu_output)
//   --export-interval N  export (and probe node 96) every N loop iterations
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
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
#if defined(USE_QUADRATIC_BASIS)
using Basis = TetrahedralBasisQuadratic<T>;
#else
using Basis = TetrahedralBasisLinear<T>;
#endif

static void dump(const std::string &path, const T *a, int n) {
  FILE *f = std::fopen(path.c_str(), "w");
  if (!f) return;
  for (int i = 0; i < n; i++) std::fprintf(f, "%.17g\n", a[i]);
  std::fclose(f);
}

int main(int argc, char *argv[]) {
  using Quadrature = TetrahedralQuadrature5pts;
  using Physics = NeohookeanPhysics<T>;
  using Analysis = FEAnalysis<T, Basis, Quadrature, Physics>;

  constexpr int dof_per_node = 3;

  bool smoke_test = false, ep = false;
  std::string filename("../input/fuselage 5086 elements.inp"), dump_prefix, out_dir;
  int export_every = 0;
  T vz = -10, location = 0.15 - 0.00005;
  double dt = 1e-7, time_end = -1;
  for (int i = 1; i < argc; i++) {
    std::string a(argv[i]);
    auto next = [&]() { return i + 1 < argc ? std::string(argv[++i]) : std::string(); };
    if (a == "-h" || a == "--help") {
      std::printf("Usage: ./impact [--smoke] [--mesh FILE] [--velocity VZ] [--wall Z] [--dt DT]"
                  " [--tend T] [--ep] [--dump PREFIX] [--out DIR]\n");
      return 0;
    } else if (a == "--smoke") {
      smoke_test = true;
    } else if (a == "--mesh") {
      filename = next();
    } else if (a == "--velocity") {
      vz = std::atof(next().c_str());
    } else if (a == "--wall") {
      location = std::atof(next().c_str());
    } else if (a == "--dt") {
      dt = std::atof(next().c_str());
    } else if (a == "--tend") {
      time_end = std::atof(next().c_str());
    } else if (a == "--ep") {
      ep = true;
    } else if (a == "--dump") {
      dump_prefix = next();
    } else if (a == "--out") {
      out_dir = next();
    } else if (a == "--export-interval") {
      export_every = std::atoi(next().c_str());
    }
  }
  if (time_end < 0) time_end = smoke_test ? dt * 10 : 0.5;

  Mesh<T, Basis::nodes_per_element> tensile;

  // Material Properties
  T E = 68.9E9;  // Pa
  T rho = 2700;  // kg/m3
  T nu = 0.33;
  T beta = 0.0;
  T H = 10;
  T Y0 = 1.9 * std::sqrt(3.0);
  std::string name = "AL6061";

  Elastoplastic<T, dof_per_node> material(E, rho, nu, beta, H, Y0, name);
  std::printf("Material: %s\n", name.c_str());
  int load_success = tensile.load_mesh(filename);
  if (load_success != 0) {
    std::cerr << "Error loading mesh" << std::endl;
    return 1;
  }

  T init_position[] = {0, 0, 0};
  T init_velocity[] = {0, 0.0, vz};

  const int normal = 1;
  std::string wall_name = "Wall";

  int export_interval = INT_MAX;
#ifdef CPPIMPACT_DEBUG_MODE
  export_interval = 100;
#endif
  if (export_every > 0) export_interval = export_every;

  Wall<T, 2, Basis> w(wall_name, location, E * 10, tensile.slave_nodes, tensile.num_slave_nodes,
                      normal);

  Dynamics<T, Basis, Analysis, Quadrature> dyna(&tensile, &material, &w);
  if (ep) dyna.options().physics = CPI_PHYSICS_EP;
  if (!out_dir.empty()) dyna.set_export_directory(out_dir);
  dyna.initialize(init_position, init_velocity);

  auto start = std::chrono::high_resolution_clock::now();
  dyna.solve(dt, time_end, export_interval);
  auto end = std::chrono::high_resolution_clock::now();
  std::chrono::duration<double> elapsed = end - start;
  std::cout << "Elapsed time: " << elapsed.count() << " seconds" << std::endl;
  std::printf("Steps: %d\n", dyna.timestep);

  if (!dump_prefix.empty()) {
    dump(dump_prefix + "_x.txt", dyna.global_xloc, dyna.ndof);
    dump(dump_prefix + "_v.txt", dyna.vel, dyna.ndof);
    dump(dump_prefix + "_vi.txt", dyna.vel_i, dyna.ndof);
    dump(dump_prefix + "_m.txt", dyna.global_mass, dyna.ndof);
  }
  return 0;
}
