// Mesh<T, nodes_per_element> with the reference's public surface
// (mesh.h:11-55): the same members, 0-based element_nodes [E][npe] and
// xloc [N][3].  load_mesh reads the same .inp subset (mesh.h:57-267) through
// cpi_inp_read and returns 0 on success, 1 when the file cannot be opened.
#pragma once

#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "cppimpact_defs.h"

template <typename T, int nodes_per_element>
class Mesh {
 public:
  int num_elements, num_nodes, num_node_sets, num_fixed_nodes, num_slave_nodes;
  int *element_nodes, *node_set_starts, *node_set_indices, *fixed_nodes, *slave_nodes;
  T *xloc;
  std::vector<std::string> node_set_names;

  Mesh()
      : num_elements(0),
        num_nodes(0),
        num_node_sets(0),
        num_fixed_nodes(0),
        num_slave_nodes(0),
        element_nodes(nullptr),
        node_set_starts(nullptr),
        node_set_indices(nullptr),
        fixed_nodes(nullptr),
        slave_nodes(nullptr),
        xloc(nullptr) {}

  ~Mesh() { release(); }

  Mesh(const Mesh &) = delete;
  Mesh &operator=(const Mesh &) = delete;

  int load_mesh(std::string filename) {
    static_assert(nodes_per_element == 4 || nodes_per_element == 10,
                  "C3D4 or C3D10 meshes only");
    cpi_inp_counts c{};
    int rc = cpi_inp_read(filename.c_str(), nodes_per_element, &c, nullptr, nullptr, nullptr,
                          nullptr, nullptr, nullptr, nullptr);
    if (rc == CPI_ERR_IO && std::strstr(cpi_last_error(nullptr), "Failed to open")) {
      std::cerr << "Failed to open file: " << filename << std::endl;
      return 1;
    }
    cpiErrchk(rc, nullptr);
    release();
    num_nodes = c.num_nodes;
    num_elements = c.num_elements;
    num_node_sets = c.num_node_sets;
    num_fixed_nodes = c.num_fixed_nodes;
    num_slave_nodes = c.num_slave_nodes;
    xloc = new T[3 * (size_t)num_nodes];
    element_nodes = new int[(size_t)nodes_per_element * num_elements];
    node_set_starts = new int[num_node_sets + 1];
    node_set_indices = new int[c.num_set_indices > 0 ? c.num_set_indices : 1];
    fixed_nodes = new int[num_fixed_nodes > 0 ? num_fixed_nodes : 1];
    slave_nodes = new int[num_slave_nodes > 0 ? num_slave_nodes : 1];
    std::vector<double> x(3 * (size_t)num_nodes);
    std::vector<char> names(c.names_bytes > 0 ? c.names_bytes : 1);
    rc = cpi_inp_read(filename.c_str(), nodes_per_element, &c, x.data(), element_nodes,
                      node_set_starts, node_set_indices, fixed_nodes, slave_nodes,
                      names.data());
    cpiErrchk(rc, nullptr);
    for (size_t i = 0; i < x.size(); i++) xloc[i] = T(x[i]);
    node_set_names.clear();
    const char *p = names.data();
    for (int s = 0; s < num_node_sets; s++) {
      node_set_names.emplace_back(p);
      p += node_set_names.back().size() + 1;
    }
    return 0;
  }

 private:
  void release() {
    delete[] element_nodes;
    delete[] node_set_starts;
    delete[] node_set_indices;
    delete[] fixed_nodes;
    delete[] slave_nodes;
    delete[] xloc;
    element_nodes = node_set_starts = node_set_indices = fixed_nodes = slave_nodes = nullptr;
    xloc = nullptr;
  }
};
