// Wall<T, dim, Basis> (wall.h:6-45): same members and constructor; the
// penalty itself is applied on the GPU (a_dim += -d * normal * k / m with
// d = (x_dim + du_dim - location) * normal < 0); detect_contact is the same
// rule on host arrays.
#pragma once

#include <string>

#include "cppimpact_defs.h"

template <typename T, int dim, class Basis>
class Wall {
 public:
  std::string name;
  T location;
  T stiffness;
  int *slave_node_indices;
  int num_slave_nodes;
  int normal;
  T norm_stiffness;

  Wall(std::string name, T location, T stiffness, int *slave_node_indices, int num_slave_nodes,
       int normal)
      : name(name),
        location(location),
        stiffness(stiffness),
        slave_node_indices(slave_node_indices),
        num_slave_nodes(num_slave_nodes),
        normal(normal) {
    norm_stiffness = normal * stiffness;
  }

  ~Wall() {}

  void detect_contact(T *global_acc, int node_idx, T *node_pos, T *node_mass) {
    T wall_distance = (node_pos[dim] - location) * normal;
    if (wall_distance < 0.0)
      global_acc[3 * node_idx + dim] += -1 * (1 / node_mass[dim]) * wall_distance * norm_stiffness;
  }
};
