// Dynamics<T, Basis, Analysis, Quadrature> for the HIP backend.
//
// The public surface of the reference's CPU and CUDA Dynamics
// (dynamics.h:16-513, dynamics.cuh:16-525): constructor, initialize, solve,
// debug_strain, export_to_vtk, probe_node, probe_element, and the public
// host state arrays.  Everything here is host C++; solve() and
// debug_strain() run on the GPU through the C ABI of libcppimpact_hip.so
// (include/cppimpact_hip.h).  A driver selects this backend the way it
// selects the CUDA one (test.cu:14-18):
//
//   #if defined(CPPIMPACT_HIP_BACKEND)
//   #include "dynamics_hip.h"
//   #elif defined(CPPIMPACT_CUDA_BACKEND) ...
//
// Behaviour kept from the reference: the loop runs while the accumulated
// time <= time_end (dynamics.h:471,509); iteration k exports when
// k % export_interval == 0 (so k = 0 always does); after solve() returns the
// host arrays hold the final state, as the CPU path leaves them.  Changed:
// timestep starts at 0 (uninitialised in the reference), probes print the
// simulation time, no per-step printf, tet10 cells are VTK type 24, and
// failures of the library print its message with file/line and exit, as
// gpuErrchk does.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

#include "basematerial.h"
#include "elastoplastic.h"
#include "mesh.h"
#include "tetrahedral.h"
#include "wall.h"

template <typename T, class Basis, class Analysis, class Quadrature>
class Dynamics {
  static_assert(std::is_same<T, double>::value, "the host state of the HIP backend is double");

 private:
  cpi_options opts_;
  std::string export_dir_ = "../gpu_output";
  double time_ = 0.0;

  void probe_node_data(int node_id, std::ostringstream &stream) {
    if (node_id < 0 || node_id >= mesh->num_nodes) {
      stream << "Node ID out of range.\n";
      return;
    }
    const int i = 3 * node_id;
    stream << "  Position: (" << global_xloc[i] << ", " << global_xloc[i + 1] << ", "
           << global_xloc[i + 2] << ")\n"
           << "  Velocity: (" << vel[i] << ", " << vel[i + 1] << ", " << vel[i + 2] << ")\n"
           << "  Acceleration: (" << global_acc[i] << ", " << global_acc[i + 1] << ", "
           << global_acc[i + 2] << ")\n"
           << "  Mass: (" << global_mass[i] << ", " << global_mass[i + 1] << ", "
           << global_mass[i + 2] << ")\n";
  }

  cpi_ctx *create_context() {
    cpi_mesh_desc d{};
    d.num_nodes = mesh->num_nodes;
    d.num_elements = mesh->num_elements;
    d.nodes_per_element = nodes_per_element;
    d.xloc = mesh->xloc;
    d.element_nodes = mesh->element_nodes;
    if (wall) {
      d.num_slave_nodes = wall->num_slave_nodes;
      d.slave_nodes = wall->slave_node_indices;
    }
    d.num_fixed_nodes = mesh->num_fixed_nodes;
    d.fixed_nodes = mesh->fixed_nodes;
    cpi_material m{};
    m.E = material->E;
    m.rho = material->rho;
    m.nu = material->nu;
    if (auto *ep = dynamic_cast<const Elastoplastic<T, dof_per_node> *>(material)) {
      m.beta = ep->get_beta();
      m.H = ep->get_H();
      m.Y0 = ep->get_Y0();
    }
    cpi_wall w{};
    if (wall) {
      w.enabled = 1;
      w.dim = 2;
      w.normal = wall->normal;
      w.location = wall->location;
      w.stiffness = wall->stiffness;
    }
    cpi_ctx *ctx = nullptr;
    cpiErrchk(cpi_create(&d, &m, wall ? &w : nullptr, &opts_, &ctx), nullptr);
    return ctx;
  }

  // device state -> the public host arrays
  void fetch(cpi_ctx *ctx) {
    std::vector<double> m(mesh->num_nodes);
    cpiErrchk(cpi_get_state(ctx, global_xloc, vel, global_acc, m.data(), vel_i), ctx);
    for (int i = 0; i < mesh->num_nodes; i++)
      global_mass[3 * i] = global_mass[3 * i + 1] = global_mass[3 * i + 2] = m[i];
    cpiErrchk(cpi_get_increment(ctx, global_dof), ctx);
    cpiErrchk(cpi_get_stress_strain(ctx, global_strains, global_stress), ctx);
  }

 public:
  int *reduced_nodes;
  int reduced_dofs_size;
  int ndof;
  static constexpr int nodes_per_element = Basis::nodes_per_element;
  static constexpr int spatial_dim = Basis::spatial_dim;
  static constexpr int dof_per_node = spatial_dim;
  static constexpr int num_quadrature_pts = Quadrature::num_quadrature_pts;

  Mesh<T, nodes_per_element> *mesh;
  BaseMaterial<T, dof_per_node> *material;
  Wall<T, 2, Basis> *wall;
  T *global_xloc;
  T *vel;
  T *global_strains;
  T *global_stress;
  T *global_dof;
  T *global_acc;
  T *global_mass;
  T *vel_i;
  int timestep;

  Dynamics(Mesh<T, nodes_per_element> *input_mesh, BaseMaterial<T, dof_per_node> *input_material,
           Wall<T, 2, Basis> *input_wall = nullptr)
      : reduced_nodes(nullptr),
        reduced_dofs_size(0),
        mesh(input_mesh),
        material(input_material),
        wall(input_wall),
        global_xloc(new T[mesh->num_nodes * dof_per_node]()),
        vel(new T[mesh->num_nodes * dof_per_node]()),
        global_strains(new T[mesh->num_nodes * 6]()),
        global_stress(new T[mesh->num_nodes * 6]()),
        global_dof(new T[mesh->num_nodes * dof_per_node]()),
        global_acc(new T[mesh->num_nodes * dof_per_node]()),
        global_mass(new T[mesh->num_nodes * dof_per_node]()),
        vel_i(new T[mesh->num_nodes * dof_per_node]()),
        timestep(0) {
    ndof = mesh->num_nodes * dof_per_node;
    cpi_default_options(&opts_);
    opts_.quadrature = Quadrature::cpi_rule;
  }

  ~Dynamics() {
    delete[] reduced_nodes;
    delete[] vel;
    delete[] global_xloc;
    delete[] global_strains;
    delete[] global_stress;
    delete[] global_dof;
    delete[] global_acc;
    delete[] global_mass;
    delete[] vel_i;
  }

  Dynamics(const Dynamics &) = delete;
  Dynamics &operator=(const Dynamics &) = delete;

  // Backend options (force model, precision, scatter, surrogate, ...); the
  // defaults reproduce the reference's step.
  cpi_options &options() { return opts_; }
  void set_export_directory(const std::string &dir) { export_dir_ = dir; }

  // dynamics.h:99-111
  void initialize(T init_position[dof_per_node], T init_velocity[dof_per_node]) {
    std::cout << "ndof: " << ndof << std::endl;
    for (int i = 0; i < mesh->num_nodes; i++) {
      for (int k = 0; k < 3; k++) {
        vel[3 * i + k] = init_velocity[k];
        mesh->xloc[3 * i + k] = mesh->xloc[3 * i + k] + init_position[k];
      }
    }
  }

  // dynamics.h:113-254: legacy ASCII VTK, same fields and order
  void export_to_vtk(int timestep, T *vel_i, T *acc_i, T *mass_i) {
    const std::string filename = export_dir_ + "/simulation_" + std::to_string(timestep) + ".vtk";
    std::ofstream f(filename);
    if (!f.is_open()) {
      std::cerr << "Failed to open " << filename << std::endl;
      return;
    }
    f << "# vtk DataFile Version 3.0\n";
    f << "FEA simulation data\n";
    f << "ASCII\n";
    f << "DATASET UNSTRUCTURED_GRID\n";
    const double threshold = 1e15;
    f << "POINTS " << mesh->num_nodes << " float\n";
    f << std::fixed << std::setprecision(6);
    for (int i = 0; i < mesh->num_nodes; ++i) {
      for (int k = 0; k < 3; k++) {
        T v = global_xloc[3 * i + k];
        if (std::isnan(v) || std::isinf(v) || std::abs(v) > threshold) {
          std::printf("Invalid value detected in %c-coordinate at node %d: %f, setting to 0.\n",
                      "xyz"[k], i, v);
          v = 0.0;
        }
        f << v << (k < 2 ? " " : "\n");
      }
    }
    f << "CELLS " << mesh->num_elements << " " << mesh->num_elements * (nodes_per_element + 1)
      << "\n";
    for (int i = 0; i < mesh->num_elements; ++i) {
      f << nodes_per_element;
      for (int j = 0; j < nodes_per_element; ++j)
        f << " " << mesh->element_nodes[nodes_per_element * i + j];
      f << "\n";
    }
    f << "CELL_TYPES " << mesh->num_elements << "\n";
    for (int i = 0; i < mesh->num_elements; ++i) f << (nodes_per_element == 10 ? "24\n" : "10\n");
    f << "POINT_DATA " << mesh->num_nodes << "\n";
    auto vec3 = [&](const char *name, const T *a, int stride, int off, bool nan_zero) {
      f << "VECTORS " << name << " double\n";
      for (int i = 0; i < mesh->num_nodes; ++i)
        for (int j = 0; j < 3; ++j) {
          T v = a[stride * i + off + j];
          if (nan_zero && std::isnan(v)) {
            std::cerr << "NaN detected in " << name << " at node " << i << ", component " << j
                      << std::endl;
            v = 0.0;
          }
          f << v << (j < 2 ? " " : "\n");
        }
    };
    vec3("velocity", vel_i, 3, 0, true);
    vec3("strain1", global_strains, 6, 0, false);
    vec3("strain2", global_strains, 6, 3, false);
    vec3("stress1", global_stress, 6, 0, false);
    vec3("stress2", global_stress, 6, 3, false);
    vec3("acceleration", acc_i, 3, 0, true);
    f << "VECTORS mass double\n";
    for (int i = 0; i < mesh->num_nodes; ++i)
      for (int j = 0; j < 3; ++j) {
        T v = mass_i[3 * i + j];
        if (std::isnan(v) || v < 0.0)
          std::cerr << "Invalid value detected in mass at node " << i << ", component " << j
                    << std::endl;
        f << v << (j < 2 ? " " : "\n");
      }
    std::cout << "Exported " << filename << std::endl;
  }

  // dynamics.h:256-274
  void probe_node(int node_id) {
    std::string filename = "../output/nodes/node_" + std::to_string(node_id) + ".txt";
    if (timestep == 0) std::remove(filename.c_str());
    std::ofstream file(filename, std::ios::app);
    std::ostringstream node_data;
    probe_node_data(node_id, node_data);
    file << "Timestep " << timestep << ", Time: " << std::fixed << std::setprecision(2) << time_
         << "s:\n";
    file << node_data.str() << "\n";
  }

  // dynamics.h:276-301
  void probe_element(int element_id) {
    if (element_id < 0 || element_id >= mesh->num_elements) {
      std::cerr << "Element ID out of range.\n";
      return;
    }
    std::string filename = "../output/elements/element_" + std::to_string(element_id) + ".txt";
    if (timestep == 0) std::remove(filename.c_str());
    std::ofstream file(filename, std::ios::app);
    const int *nodes = &mesh->element_nodes[nodes_per_element * element_id];
    file << "Timestep " << timestep << ", Time: " << std::fixed << std::setprecision(2) << time_
         << "s:\n"
         << "Element " << element_id << " consists of nodes:\n";
    for (int i = 0; i < nodes_per_element; ++i) {
      std::ostringstream node_data;
      probe_node_data(nodes[i], node_data);
      file << " Node " << nodes[i] << " details:\n" << node_data.str();
    }
    file << "\n";
  }

  // dynamics.h:303-412: prescribed displacement field, total strain energy
  // and volume (on the GPU), then the deformed configuration exported
  void debug_strain(const T alpha, const int def_case) {
    std::memcpy(global_xloc, mesh->xloc, ndof * sizeof(T));
    std::memset(global_dof, 0, sizeof(T) * ndof);
    for (int i = 0; i < mesh->num_nodes; i++) {
      const T x = global_xloc[3 * i];
      if (def_case == 0) {
        if (i == 0) std::printf("Constant displacement case\n");
        global_dof[3 * i + 0] = alpha * x;
        global_dof[3 * i + 1] = -alpha * x * material->nu;
        global_dof[3 * i + 2] = -alpha * x * material->nu;
      } else if (def_case == 1) {
        if (i == 0) std::printf("Linear displacement case\n");
        global_dof[3 * i + 0] = alpha * 0.5 * x * x;
        global_dof[3 * i + 1] = -alpha * 0.5 * x * x * material->nu;
        global_dof[3 * i + 2] = -alpha * 0.5 * x * x * material->nu;
      }
    }
    cpi_ctx *ctx = create_context();
    double W = 0.0, V = 0.0;
    cpiErrchk(cpi_debug_strain(ctx, alpha, def_case, &W, &V), ctx);
    cpi_destroy(ctx);
    std::printf("Total Strain Energy = %f\n", W);
    std::printf("Total Volume = %f\n", V);
    for (int i = 0; i < ndof; i++) global_xloc[i] += global_dof[i];
    export_to_vtk(0, vel, global_acc, global_mass);
  }

  // dynamics.h:414-512 on the GPU: mass, a0 = contact + gravity and the
  // half-kick in cpi_init, then one cpi_step per run of loop iterations
  // between exports
  void solve(double dt, double time_end, int export_interval) {
    std::printf("Solving dynamics\n");
    std::memcpy(global_xloc, mesh->xloc, ndof * sizeof(T));
    for (int i = 0; i < ndof; i++) global_dof[i] = global_acc[i] = global_mass[i] = vel_i[i] = 0.0;
    cpi_ctx *ctx = create_context();
    cpiErrchk(cpi_set_state(ctx, mesh->xloc, vel), ctx);
    cpiErrchk(cpi_init(ctx, dt), ctx);
    long long n = 0;
    for (double t = 0.0; t <= time_end; t += dt) n++;
    long long done = 0;
    time_ = 0.0;
    while (done < n) {
      long long chunk = n - done;
      if (export_interval > 0) {
        const long long ei = export_interval, k = timestep;
        const long long next = ((k + ei - 1) / ei) * ei;  // next exporting iteration
        chunk = std::min(chunk, next - k + 1);
      }
      chunk = std::min<long long>(chunk, 1 << 30);
      cpiErrchk(cpi_step(ctx, (int32_t)chunk), ctx);
      done += chunk;
      // iteration k = timestep + chunk - 1 just ended: as in the reference,
      // it exports and probes under its own index (dynamics.h:505-508), so
      // a probe file starts afresh at iteration 0; the probes' time is the
      // simulated time of the state they print, (k + 1) dt
      timestep += (int)chunk - 1;
      time_ += chunk * dt;
      if (export_interval > 0 && timestep % export_interval == 0) {
        fetch(ctx);
        export_to_vtk(timestep, vel_i, global_acc, global_mass);
        probe_node(96);
      }
      timestep += 1;
    }
    fetch(ctx);
    cpi_destroy(ctx);
  }
};
