// Internal declarations shared by the host-side C++ and the HIP translation
// unit of libcppimpact_hip.so.  Not part of the public ABI.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace cpi {

// Greedy colouring in element order, smallest free colour.  Returns the
// number of colours.
int greedy_coloring(int ne, int npe, const int32_t *conn, int nn,
                    int32_t *color);

// Workgroup partition for the LDS-assembled scatter.
//
// Elements are sorted along a 3-D Hilbert curve of their centroids and cut
// into blocks of at most max_elems elements and max_nodes distinct nodes.
// A node referenced by one block only is "interior" and is updated by that
// block; the others are "boundary" nodes whose force arrives as one partial
// per referencing block and is summed by the boundary-node kernel.
//
// Device node order: interior nodes of block 0, of block 1, ..., then the
// boundary nodes ordered by (first referencing block, user id).
// Device element order: block order, Hilbert order inside a block.
struct BlockPlan {
  int nblocks = 0;
  int max_local_nodes = 0;  // over all blocks (ml)
  int max_block_elems = 0;
  int kmax = 0;             // most contributions of one local node
  int n_interior = 0;       // device ids [0, n_interior) are interior nodes
  int n_boundary = 0;       // device ids [n_interior, n_interior+n_boundary)
  int n_slots = 0;
  bool grown = false;       // grown blocks (else the cut Hilbert order)
  std::vector<int32_t> elem_perm;  // device element -> user element
  std::vector<int32_t> node_perm;  // device node -> user node
  std::vector<int32_t> node_inv;   // user node -> device node
  // per block, 8 int32: elem_start, n_elem, int_start, n_int, bnd_off,
  // n_bnd, corner_off, kmax of the block
  std::vector<int32_t> meta;
  std::vector<uint16_t> conn_local;  // [E][npe] local node index
  std::vector<int32_t> bnd_nodes;    // per block boundary nodes (device id)
  std::vector<int32_t> bnd_slot;     // parallel to bnd_nodes: partial slot
  // Corner forces of a block are stored jagged-diagonal (local nodes sorted
  // by contribution count): the k-th contribution of local node l, in
  // element order, is at off[k] + l; corner j of element t writes to
  // csr_pos[t*npe + j].
  std::vector<uint16_t> csr_pos;     // [E][npe] position of each corner force
  // The same corner in 16 bits, l | k << 9 (k: the element's rank among
  // local node l's contributions, so csr_pos = off[k] + l), for the
  // per-point tet10 block kernel; exact when max_local_nodes <= 512 and
  // kmax <= 128 (corner_fits).
  std::vector<uint16_t> corner;
  bool corner_fits() const { return max_local_nodes <= 512 && kmax <= 128; }
  // fixed-stride per-block tables, local node l of block b at b*ml + l:
  std::vector<int32_t> nid_pad;      // device node id, -1 past the block's nodes
  std::vector<int32_t> slot_pad;     // partial slot, -1 for interior nodes
  std::vector<uint16_t> deg_pad;     // contribution count
  std::vector<uint16_t> off_pad;     // per block kmax+1 offsets off[k]
  std::vector<int32_t> slot_ptr;     // per boundary node, n_boundary+1
};

bool build_block_plan(int ne, int npe, const int32_t *conn, int nn,
                      const double *xloc, int max_elems, int max_nodes,
                      BlockPlan &plan, std::string &err);

// Element order sorted by colour (stable), and colour offsets.
void color_order(int ne, const int32_t *color, int ncol,
                 std::vector<int32_t> &perm, std::vector<int32_t> &offsets);

// Hilbert index of a point with 21-bit integer coordinates.
uint64_t hilbert3(uint32_t x, uint32_t y, uint32_t z);

}  // namespace cpi
