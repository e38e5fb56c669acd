// Invariants of the workgroup block plan (cpi::build_block_plan), checked on
// the host: every LDS force slot and every partial-force slot a kernel reads
// is written exactly once per step, by the right corner or block, and every
// table entry a kernel indexes with is in range.  Built by
// tests/test_plan_invariants.py with AddressSanitizer and
// UndefinedBehaviorSanitizer against cppimpact_amd/csrc/host.cpp.
//
// usage: plan_invariants  (runs the built-in meshes; exit status 0 = pass)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../cppimpact_amd/csrc/internal.h"
#include "../include/cppimpact_hip.h"

using namespace cpi;

static int failures = 0;

#define REQUIRE(cond, ...)                              \
  do {                                                  \
    if (!(cond)) {                                      \
      std::fprintf(stderr, "FAIL %s: ", label.c_str()); \
      std::fprintf(stderr, __VA_ARGS__);                \
      std::fprintf(stderr, "\n");                       \
      if (++failures > 20) std::exit(1);                \
      return;                                           \
    }                                                   \
  } while (0)

static void check(const std::string &label, int ne, int npe, const std::vector<int32_t> &conn,
                  int nn, const std::vector<double> &x, int max_elems, int max_nodes) {
  BlockPlan P;
  std::string err;
  REQUIRE(build_block_plan(ne, npe, conn.data(), nn, x.data(), max_elems, max_nodes, P, err),
          "build_block_plan: %s", err.c_str());
  const int nb = P.nblocks, ml = P.max_local_nodes, ko = P.kmax + 1;
  REQUIRE(nb > 0 && ml > 0 && ml <= max_nodes && P.max_block_elems <= max_elems,
          "sizes nb %d ml %d max_el %d", nb, ml, P.max_block_elems);
  REQUIRE(P.n_interior + P.n_boundary == nn, "node split %d + %d != %d", P.n_interior,
          P.n_boundary, nn);
  // permutations
  std::vector<int> seen(std::max(ne, nn), 0);
  for (int i = 0; i < nn; i++) {
    const int g = P.node_perm[i];
    REQUIRE(g >= 0 && g < nn && !seen[g]++, "node_perm not a permutation at %d", i);
    REQUIRE(P.node_inv[g] == i, "node_inv(node_perm(%d)) != %d", i, i);
  }
  std::fill(seen.begin(), seen.end(), 0);
  for (int e = 0; e < ne; e++) {
    const int u = P.elem_perm[e];
    REQUIRE(u >= 0 && u < ne && !seen[u]++, "elem_perm not a permutation at %d", e);
  }
  // blocks cover the elements contiguously
  int next_e = 0;
  std::vector<int> blocks_of_node(nn, 0);
  std::vector<int> slot_writes(P.n_slots, 0), slot_node(P.n_slots, -1);
  for (int b = 0; b < nb; b++) {
    const int32_t *m = &P.meta[(size_t)b * 8];
    const int es = m[0], nel = m[1];
    REQUIRE(es == next_e && nel > 0 && nel <= max_elems, "block %d elements [%d,+%d)", b, es,
            nel);
    next_e += nel;
    // local nodes: distinct device ids, padding -1 after the block's nodes
    int nl = 0;
    while (nl < ml && P.nid_pad[(size_t)b * ml + nl] >= 0) nl++;
    for (int l = nl; l < ml; l++)
      REQUIRE(P.nid_pad[(size_t)b * ml + l] == -1 && P.slot_pad[(size_t)b * ml + l] == -1 &&
                  P.deg_pad[(size_t)b * ml + l] == 0,
              "block %d padding at %d", b, l);
    REQUIRE(nl <= max_nodes, "block %d holds %d nodes", b, nl);
    std::vector<int> local_of(nn, -1);
    for (int l = 0; l < nl; l++) {
      const int d = P.nid_pad[(size_t)b * ml + l];
      REQUIRE(d < nn && local_of[d] < 0, "block %d local %d device id %d", b, l, d);
      local_of[d] = l;
      blocks_of_node[d]++;
    }
    // connectivity through the local table, corner counts, LDS force slots
    std::vector<int> deg(nl, 0);
    const int km = m[7];
    REQUIRE(km <= P.kmax, "block %d kmax %d > %d", b, km, P.kmax);
    const uint16_t *off = &P.off_pad[(size_t)b * ko];
    for (int k = 0; k < km; k++) REQUIRE(off[k] <= off[k + 1], "block %d off not monotone", b);
    const int nslots_lds = off[km];
    REQUIRE(nslots_lds == nel * npe, "block %d: %d LDS force slots for %d corners", b,
            nslots_lds, nel * npe);
    std::vector<int> pos_seen(nslots_lds, 0);
    for (int t = 0; t < nel; t++) {
      const int e = es + t, u = P.elem_perm[e];
      for (int k = 0; k < npe; k++) {
        const int l = P.conn_local[(size_t)e * npe + k];
        REQUIRE(l < nl, "element %d corner %d local id %d >= %d", e, k, l, nl);
        const int d = P.nid_pad[(size_t)b * ml + l];
        REQUIRE(d == P.node_inv[conn[(size_t)u * npe + k]],
                "element %d corner %d maps to the wrong node", e, k);
        // the r-th contribution of local node l (element order) sits at
        // off[r] + l, and no two corners share a slot
        const int r = deg[l]++;
        const int pos = P.csr_pos[(size_t)e * npe + k];
        REQUIRE(r < km && pos == off[r] + l, "element %d corner %d slot %d != off[%d] + %d", e,
                k, pos, r, l);
        REQUIRE(pos < nslots_lds && !pos_seen[pos]++, "LDS slot %d written twice", pos);
        // the 16-bit corner word: local node and rank, the same slot
        if (P.corner_fits()) {
          const int w = P.corner[(size_t)e * npe + k];
          REQUIRE((w & 511) == l && (w >> 9) == r && off[w >> 9] + (w & 511) == pos,
                  "element %d corner %d word %d != %d | %d << 9", e, k, w, l, r);
        }
      }
    }
    for (int l = 0; l < nl; l++) {
      REQUIRE(deg[l] == P.deg_pad[(size_t)b * ml + l], "block %d node %d degree %d != %d", b,
              l, deg[l], P.deg_pad[(size_t)b * ml + l]);
      // every slot the node sum reads (off[k] + l, k < deg) is one of the
      // block's corner slots: jagged-diagonal ranges cover exactly the nodes
      // with more than k contributions
      for (int k = 0; k < deg[l]; k++)
        REQUIRE(l < off[k + 1] - off[k], "block %d node %d reads past column %d", b, l, k);
      const int d = P.nid_pad[(size_t)b * ml + l], s = P.slot_pad[(size_t)b * ml + l];
      if (d < P.n_interior) {
        REQUIRE(s == -1, "interior node %d has partial slot %d", d, s);
      } else {
        REQUIRE(s >= 0 && s < P.n_slots, "boundary node %d slot %d", d, s);
        slot_writes[s]++;
        slot_node[s] = d;
      }
    }
    for (int p = 0; p < nslots_lds; p++) REQUIRE(pos_seen[p] == 1, "LDS slot %d never written", p);
  }
  REQUIRE(next_e == ne, "blocks cover %d of %d elements", next_e, ne);
  // partial slots: each written by exactly one block per step, and the
  // boundary-node kernel reads, for node n_interior + i, exactly the slots
  // written for that node
  for (int s = 0; s < P.n_slots; s++) REQUIRE(slot_writes[s] == 1, "slot %d written %d times", s,
                                              slot_writes[s]);
  REQUIRE(P.slot_ptr[0] == 0 && P.slot_ptr[P.n_boundary] == P.n_slots, "slot_ptr ends");
  for (int i = 0; i < P.n_boundary; i++) {
    const int d = P.n_interior + i;
    REQUIRE(P.slot_ptr[i] <= P.slot_ptr[i + 1], "slot_ptr not monotone at %d", i);
    REQUIRE(P.slot_ptr[i + 1] - P.slot_ptr[i] == blocks_of_node[d],
            "node %d: %d slots for %d blocks", d, P.slot_ptr[i + 1] - P.slot_ptr[i],
            blocks_of_node[d]);
    REQUIRE(blocks_of_node[d] != 1, "boundary node %d in one block only", d);
    for (int s = P.slot_ptr[i]; s < P.slot_ptr[i + 1]; s++)
      REQUIRE(slot_node[s] == d, "slot %d of node %d written for node %d", s, d, slot_node[s]);
  }
  for (int d = 0; d < P.n_interior; d++)
    REQUIRE(blocks_of_node[d] == 1, "interior node %d in %d blocks", d, blocks_of_node[d]);
  std::printf("ok %s: %d elements, %d %s blocks, %d interior + %d boundary nodes, %d slots, "
              "kmax %d, corner words %s\n",
              label.c_str(), ne, nb, P.grown ? "grown" : "Hilbert", P.n_interior, P.n_boundary,
              P.n_slots, P.kmax, P.corner_fits() ? "fit" : "do not fit");
}

// a fan of 2m tet4 around node 0: ring nodes 1..m on the unit circle,
// apexes m+1 (z = 1) and m+2 (z = -1); node 0 is in every element
static void fan(const std::string &label, int m, int max_el, bool fits) {
  const int nn = m + 3, ne = 2 * m;
  std::vector<double> x(3 * (size_t)nn, 0.0);
  for (int i = 0; i < m; i++) {
    x[3 * (1 + i)] = std::cos(2.0 * M_PI * i / m);
    x[3 * (1 + i) + 1] = std::sin(2.0 * M_PI * i / m);
  }
  x[3 * (m + 1) + 2] = 1.0;
  x[3 * (m + 2) + 2] = -1.0;
  std::vector<int32_t> c;
  for (int i = 0; i < m; i++) {
    const int a = 1 + i, b = 1 + (i + 1) % m;
    c.insert(c.end(), {0, a, b, m + 1});
    c.insert(c.end(), {0, b, a, m + 2});
  }
  BlockPlan P;
  std::string err;
  if (!build_block_plan(ne, 4, c.data(), nn, x.data(), max_el, 384, P, err)) {
    std::fprintf(stderr, "FAIL %s: %s\n", label.c_str(), err.c_str());
    failures++;
    return;
  }
  if (P.corner_fits() != fits) {
    std::fprintf(stderr, "FAIL %s: kmax %d, corner words %s\n", label.c_str(), P.kmax,
                 P.corner_fits() ? "fit" : "do not fit");
    failures++;
  }
  check(label, ne, 4, c, nn, x, max_el, 384);
}

static void box(const std::string &label, int n, int quad, double jitter, int max_el,
                int max_nodes) {
  cpi_box_spec s{};
  s.nx = s.ny = s.nz = n;
  s.quadratic = quad;
  s.lx = s.ly = s.lz = 0.25;
  s.jitter = jitter;
  s.seed = 7;
  int32_t nn = 0, ne = 0;
  if (cpi_mesh_box(&s, &nn, &ne, nullptr, nullptr) != 0) std::exit(2);
  const int npe = quad ? 10 : 4;
  std::vector<double> x(3 * (size_t)nn);
  std::vector<int32_t> c((size_t)npe * ne);
  if (cpi_mesh_box(&s, &nn, &ne, x.data(), c.data()) != 0) std::exit(2);
  check(label, ne, npe, c, nn, x, max_el, max_nodes);
}

static void cylinder(const std::string &label, int nt, int nz, int nr, int max_el,
                     int max_nodes) {
  cpi_cylinder_spec s{};
  s.ntheta = nt;
  s.nz = nz;
  s.nr = nr;
  s.quadratic = 1;
  s.radius = 1.0;
  s.length = 0.5;
  s.thickness = 0.02;
  int32_t nn = 0, ne = 0;
  if (cpi_mesh_cylinder(&s, &nn, &ne, nullptr, nullptr) != 0) std::exit(2);
  std::vector<double> x(3 * (size_t)nn);
  std::vector<int32_t> c(10 * (size_t)ne);
  if (cpi_mesh_cylinder(&s, &nn, &ne, x.data(), c.data()) != 0) std::exit(2);
  check(label, ne, 10, c, nn, x, max_el, max_nodes);
}

int main() {
  // the element and node limits cpi_create derives for each kernel
  box("tet4 fp32 (384, 384)", 12, 0, 0.1, 384, 384);
  box("tet4 fp64 (384, 308)", 12, 0, 0.1, 384, 308);
  box("tet4 small blocks (37, 64)", 7, 0, 0.05, 37, 64);
  box("tet10 (128, 911)", 5, 1, 0.05, 128, 911);
  box("tet10 (37, 200)", 4, 1, 0.0, 37, 200);
  box("tet4 one block", 2, 0, 0.0, 384, 384);
  cylinder("cylinder tet10 (128, 911)", 48, 6, 3, 128, 911);
  // a node in 200 elements: one 384-element block exceeds the 16-bit
  // corner words' 128 ranks, 128-element blocks fit them
  fan("tet4 fan of 200 (384, 384)", 100, 384, false);
  fan("tet4 fan of 200 (128, 384)", 100, 128, true);
  // a node of no element (left out of every block)
  {
    cpi_box_spec s{};
    s.nx = s.ny = s.nz = 3;
    s.lx = s.ly = s.lz = 0.1;
    int32_t nn = 0, ne = 0;
    cpi_mesh_box(&s, &nn, &ne, nullptr, nullptr);
    std::vector<double> x(3 * (size_t)nn + 3, 0.05);
    std::vector<int32_t> c(4 * (size_t)ne);
    cpi_mesh_box(&s, &nn, &ne, x.data(), c.data());
    check("tet4 with an unused node", ne, 4, c, nn + 1, x, 20, 40);
  }
  if (failures) {
    std::fprintf(stderr, "%d failures\n", failures);
    return 1;
  }
  std::printf("all plans hold\n");
  return 0;
}
