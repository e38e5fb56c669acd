// Host-side services of libcppimpact_hip.so that need no GPU: structured
// mesh generators, the .inp reader/writer, greedy element colouring and the
// workgroup partition used by the LDS-assembled scatter.
#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <numeric>
#include <queue>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cppimpact_hip.h"
#include "internal.h"

namespace cpi {

thread_local std::string g_error;

void set_global_error(const std::string &s) { g_error = s; }
const char *global_error() { return g_error.c_str(); }

// ---------------------------------------------------------------------------
// colouring
// ---------------------------------------------------------------------------

int greedy_coloring(int ne, int npe, const int32_t *conn, int nn,
                    int32_t *color) {
  // colour sets per node as bit masks, widened 64 colours at a time
  int words = 1;
  std::vector<uint64_t> mask((size_t)nn * words, 0);
  int ncol = 0;
  for (int e = 0; e < ne; e++) {
    const int32_t *c = conn + (size_t)npe * e;
    int found = -1;
    while (found < 0) {
      for (int w = 0; w < words && found < 0; w++) {
        uint64_t used = 0;
        for (int j = 0; j < npe; j++) used |= mask[(size_t)c[j] * words + w];
        if (~used) found = 64 * w + __builtin_ctzll(~used);
      }
      if (found >= 0) break;
      std::vector<uint64_t> m2((size_t)nn * (words + 1), 0);
      for (int i = 0; i < nn; i++)
        std::copy(mask.begin() + (size_t)i * words,
                  mask.begin() + (size_t)(i + 1) * words,
                  m2.begin() + (size_t)i * (words + 1));
      mask.swap(m2);
      words++;
    }
    color[e] = found;
    ncol = std::max(ncol, found + 1);
    for (int j = 0; j < npe; j++)
      mask[(size_t)c[j] * words + found / 64] |= uint64_t(1) << (found % 64);
  }
  return ncol;
}

void color_order(int ne, const int32_t *color, int ncol,
                 std::vector<int32_t> &perm, std::vector<int32_t> &offsets) {
  offsets.assign(ncol + 1, 0);
  for (int e = 0; e < ne; e++) offsets[color[e] + 1]++;
  for (int c = 0; c < ncol; c++) offsets[c + 1] += offsets[c];
  std::vector<int32_t> pos(offsets.begin(), offsets.end() - 1);
  perm.assign(ne, 0);
  for (int e = 0; e < ne; e++) perm[pos[color[e]]++] = e;
}

// ---------------------------------------------------------------------------
// Hilbert curve: J. Skilling, "Programming the Hilbert curve", AIP Conf.
// Proc. 707, 381 (2004) -- axes-to-transpose, then bit interleave.
// ---------------------------------------------------------------------------

uint64_t hilbert3(uint32_t x, uint32_t y, uint32_t z) {
  const int bits = 21;
  uint32_t X[3] = {x, y, z};
  const uint32_t M = 1u << (bits - 1);
  for (uint32_t Q = M; Q > 1; Q >>= 1) {
    uint32_t P = Q - 1;
    for (int i = 0; i < 3; i++) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 3; i++) X[i] ^= X[i - 1];
  uint32_t t = 0;
  for (uint32_t Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; i++) X[i] ^= t;
  uint64_t key = 0;
  for (int b = bits - 1; b >= 0; b--)
    for (int i = 0; i < 3; i++) key = (key << 1) | ((X[i] >> b) & 1u);
  return key;
}

// ---------------------------------------------------------------------------
// workgroup partition
// ---------------------------------------------------------------------------

bool build_block_plan(int ne, int npe, const int32_t *conn, int nn,
                      const double *xloc, int max_elems, int max_nodes,
                      BlockPlan &P, std::string &err) {
  if ((long)max_elems * npe > 65535) {
    err = "block size too large for 16-bit contribution indices";
    return false;
  }
  if (max_nodes < npe || max_nodes > 65535) {
    err = "invalid local node limit";
    return false;
  }
  // 1. Hilbert order of element centroids (corner nodes)
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < nn; i++)
    for (int k = 0; k < 3; k++) {
      lo[k] = std::min(lo[k], xloc[3 * i + k]);
      hi[k] = std::max(hi[k], xloc[3 * i + k]);
    }
  double span = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2], 1e-300});
  std::vector<uint64_t> key(ne);
  std::vector<double> cen(3 * (size_t)ne);
  const double scale = double((1u << 21) - 1) / span;
  for (int e = 0; e < ne; e++) {
    double c[3] = {0, 0, 0};
    for (int j = 0; j < 4; j++)
      for (int k = 0; k < 3; k++) c[k] += xloc[3 * conn[(size_t)npe * e + j] + k];
    uint32_t q[3];
    for (int k = 0; k < 3; k++) {
      cen[3 * (size_t)e + k] = 0.25 * c[k];
      double v = (0.25 * c[k] - lo[k]) * scale;
      v = std::min(std::max(v, 0.0), double((1u << 21) - 1));
      q[k] = (uint32_t)v;
    }
    key[e] = hilbert3(q[0], q[1], q[2]);
  }
  std::vector<int32_t> horder(ne);
  std::iota(horder.begin(), horder.end(), 0);
  std::stable_sort(horder.begin(), horder.end(),
                   [&](int a, int b) { return key[a] < key[b]; });

  // 2. two partitions under both limits, the one with fewer partial slots
  // (the sum over shared nodes of the blocks that hold them) kept: the
  // Hilbert order cut greedily into blocks, which gives the octree cells of
  // a regular grid, and blocks grown from seeds taken in Hilbert order,
  // each adding the candidate element that shares the most nodes with it
  // (ties: the one nearest the seed), which follows a thin shell through
  // its thickness where the Hilbert cuts split it.
  std::vector<int32_t> stamp(nn, -1);
  auto count_fresh = [&](int e, int cur) {
    const int32_t *c = conn + (size_t)npe * e;
    int fresh = 0;
    for (int j = 0; j < npe; j++) {
      bool dup = false;
      for (int k = 0; k < j; k++) dup |= (c[k] == c[j]);
      if (!dup && stamp[c[j]] != cur) fresh++;
    }
    return fresh;
  };
  auto slots_of = [&](const std::vector<int32_t> &ord, const std::vector<int32_t> &bs) {
    std::fill(stamp.begin(), stamp.end(), -1);
    std::vector<int32_t> cnt(nn, 0);
    for (int b = 0; b + 1 < (int)bs.size(); b++)
      for (int t = bs[b]; t < bs[b + 1]; t++)
        for (int j = 0; j < npe; j++) {
          const int g = conn[(size_t)npe * ord[t] + j];
          if (stamp[g] != b) {
            stamp[g] = b;
            cnt[g]++;
          }
        }
    long s = 0;
    for (int i = 0; i < nn; i++)
      if (cnt[i] > 1) s += cnt[i];
    return s;
  };
  // 2a. greedy chunking of the Hilbert order
  std::vector<int32_t> hstart;
  {
    std::fill(stamp.begin(), stamp.end(), -1);
    int cur = -1, cur_elems = 0, cur_nodes = 0;
    for (int t = 0; t < ne; t++) {
      const int e = horder[t];
      const int fresh = cur >= 0 ? count_fresh(e, cur) : 0;
      if (cur < 0 || cur_elems + 1 > max_elems || cur_nodes + fresh > max_nodes) {
        cur++;
        hstart.push_back(t);
        cur_elems = 0;
        cur_nodes = 0;
      }
      const int32_t *c = conn + (size_t)npe * e;
      for (int j = 0; j < npe; j++)
        if (stamp[c[j]] != cur) {
          stamp[c[j]] = cur;
          cur_nodes++;
        }
      cur_elems++;
    }
    hstart.push_back(ne);
  }
  // 2b. grown blocks
  std::vector<int32_t> gorder, gstart;
  {
    std::vector<int32_t> nptr(nn + 1, 0), nel;
    for (size_t k = 0; k < (size_t)ne * npe; k++) nptr[conn[k] + 1]++;
    for (int i = 0; i < nn; i++) nptr[i + 1] += nptr[i];
    nel.resize(nptr[nn]);
    {
      std::vector<int32_t> fill(nptr.begin(), nptr.end() - 1);
      for (int e = 0; e < ne; e++)
        for (int j = 0; j < npe; j++) nel[fill[conn[(size_t)npe * e + j]]++] = e;
    }
    std::fill(stamp.begin(), stamp.end(), -1);
    std::vector<char> done(ne, 0);
    std::vector<int32_t> shared(ne, 0), touched;
    gorder.reserve(ne);
    typedef std::pair<std::pair<int, double>, int> Ent;  // (shared, -dist^2), element
    int seedp = 0, cur = 0;
    while ((int)gorder.size() < ne) {
      while (done[horder[seedp]]) seedp++;
      const int seed = horder[seedp];
      const double *sc = &cen[3 * (size_t)seed];
      auto dist = [&](int f) {
        const double *fc = &cen[3 * (size_t)f];
        return (fc[0] - sc[0]) * (fc[0] - sc[0]) + (fc[1] - sc[1]) * (fc[1] - sc[1]) +
               (fc[2] - sc[2]) * (fc[2] - sc[2]);
      };
      gstart.push_back((int)gorder.size());
      std::priority_queue<Ent> pq;
      int nodes = 0, els = 0;
      touched.clear();
      auto add = [&](int e) {
        done[e] = 1;
        gorder.push_back(e);
        els++;
        for (int j = 0; j < npe; j++) {
          const int g = conn[(size_t)npe * e + j];
          if (stamp[g] == cur) continue;
          stamp[g] = cur;
          nodes++;
          for (int k = nptr[g]; k < nptr[g + 1]; k++) {
            const int f = nel[k];
            if (done[f]) continue;
            if (shared[f]++ == 0) touched.push_back(f);
            pq.push(Ent({shared[f], -dist(f)}, f));
          }
        }
      };
      add(seed);
      while (els < max_elems) {
        int pick = -1;
        while (!pq.empty()) {
          const Ent &top = pq.top();
          if (done[top.second] || top.first.first != shared[top.second]) {
            pq.pop();
            continue;
          }
          pick = top.second;
          break;
        }
        if (pick < 0) {
          // no neighbour left: the next element of the Hilbert order
          int sp = seedp;
          while (sp < ne && done[horder[sp]]) sp++;
          if (sp >= ne) break;
          pick = horder[sp];
        }
        if (nodes + count_fresh(pick, cur) > max_nodes) break;
        add(pick);
      }
      for (int f : touched) shared[f] = 0;
      cur++;
    }
    gstart.push_back(ne);
  }
  const bool grown = slots_of(gorder, gstart) < slots_of(horder, hstart);
  P.grown = grown;
  const std::vector<int32_t> &order = grown ? gorder : horder;
  std::vector<int32_t> blk_start = grown ? gstart : hstart;
  const int nb = (int)blk_start.size() - 1;

  // 3. per node: number of referencing blocks and the first one
  std::vector<int32_t> nblk(nn, 0), first(nn, -1);
  std::fill(stamp.begin(), stamp.end(), -1);
  std::vector<std::vector<int32_t>> bnodes(nb);
  for (int b = 0; b < nb; b++) {
    for (int t = blk_start[b]; t < blk_start[b + 1]; t++) {
      const int32_t *c = conn + (size_t)npe * order[t];
      for (int j = 0; j < npe; j++)
        if (stamp[c[j]] != b) {
          stamp[c[j]] = b;
          nblk[c[j]]++;
          if (first[c[j]] < 0) first[c[j]] = b;
          bnodes[b].push_back(c[j]);
        }
    }
    std::sort(bnodes[b].begin(), bnodes[b].end());
  }

  // 4. device node numbering
  P.node_perm.assign(nn, -1);
  P.node_inv.assign(nn, -1);
  int next = 0;
  std::vector<int32_t> int_start(nb), n_int(nb);
  for (int b = 0; b < nb; b++) {
    int_start[b] = next;
    for (int32_t g : bnodes[b])
      if (nblk[g] == 1) {
        P.node_inv[g] = next;
        P.node_perm[next++] = g;
      }
    n_int[b] = next - int_start[b];
  }
  P.n_interior = next;
  for (int b = 0; b < nb; b++)
    for (int32_t g : bnodes[b])
      if (nblk[g] > 1 && first[g] == b) {
        P.node_inv[g] = next;
        P.node_perm[next++] = g;
      }
  // nodes of no element go last with no partial slots: they feel only the
  // wall and gravity, as in the reference's node loop
  for (int i = 0; i < nn; i++)
    if (nblk[i] == 0) {
      P.node_inv[i] = next;
      P.node_perm[next++] = i;
    }
  P.n_boundary = next - P.n_interior;

  // 5. partial slots: boundary node b owns slots [slot_ptr[b], slot_ptr[b+1])
  P.slot_ptr.assign(P.n_boundary + 1, 0);
  for (int b = 0; b < P.n_boundary; b++)
    P.slot_ptr[b + 1] = P.slot_ptr[b] + nblk[P.node_perm[P.n_interior + b]];
  P.n_slots = P.slot_ptr[P.n_boundary];
  std::vector<int32_t> rank(P.n_boundary, 0);

  // 6. per-block local numbering.  Local nodes are sorted by their number of
  // corner references in the block (descending; ties keep interior before
  // boundary and device order).  Corner forces are stored jagged-diagonal:
  // the k-th contribution (element order) of local node l sits at
  // off[k] + l, off[k] = sum over k' < k of the count of nodes with more
  // than k' contributions; a node's sum then reads consecutive addresses
  // across the nodes of the block.
  P.nblocks = nb;
  P.meta.assign((size_t)nb * 8, 0);
  P.elem_perm.assign(ne, 0);
  P.conn_local.assign((size_t)ne * npe, 0);
  P.csr_pos.assign((size_t)ne * npe, 0);
  P.corner.assign((size_t)ne * npe, 0);
  P.bnd_nodes.clear();
  P.bnd_slot.clear();
  P.max_local_nodes = 0;
  P.max_block_elems = 0;
  P.kmax = 0;
  std::vector<int32_t> local(nn, -1);
  std::vector<std::vector<int32_t>> b_dev(nb), b_slot(nb), b_deg(nb), b_off(nb);
  for (int b = 0; b < nb; b++) {
    std::vector<int32_t> dev;  // device ids of the block's boundary nodes
    for (int32_t g : bnodes[b])
      if (nblk[g] > 1) dev.push_back(P.node_inv[g]);
    std::sort(dev.begin(), dev.end());
    // candidates in the old order: interior (device order), then boundary
    std::vector<int32_t> cand;
    for (int j = 0; j < n_int[b]; j++) cand.push_back(int_start[b] + j);
    for (int32_t d : dev) cand.push_back(d);
    const int nl = (int)cand.size();
    const int es = blk_start[b], nel = blk_start[b + 1] - blk_start[b];
    for (int j = 0; j < nl; j++) local[P.node_perm[cand[j]]] = j;
    std::vector<int32_t> deg(nl, 0);
    for (int t = 0; t < nel; t++) {
      const int32_t *c = conn + (size_t)npe * order[es + t];
      for (int j = 0; j < npe; j++) deg[local[c[j]]]++;
    }
    std::vector<int32_t> ord(nl);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int c) { return deg[a] > deg[c]; });
    std::vector<int32_t> dev_l(nl), slot_l(nl, -1), deg_l(nl);
    for (int j = 0; j < nl; j++) {
      const int o = ord[j];
      dev_l[j] = cand[o];
      deg_l[j] = deg[o];
      local[P.node_perm[cand[o]]] = j;
      if (cand[o] >= P.n_interior) {
        const int bi = cand[o] - P.n_interior;
        slot_l[j] = P.slot_ptr[bi] + rank[bi]++;
      }
    }
    const int km = nl ? deg_l[0] : 0;
    std::vector<int32_t> off(km + 1, 0);
    for (int k = 0; k < km; k++) {
      int nk = 0;
      while (nk < nl && deg_l[nk] > k) nk++;
      off[k + 1] = off[k] + nk;
    }
    int32_t *m = &P.meta[(size_t)b * 8];
    m[0] = es;
    m[1] = nel;
    m[2] = int_start[b];
    m[3] = n_int[b];
    m[4] = (int32_t)P.bnd_nodes.size();
    m[5] = (int32_t)dev.size();
    m[6] = es * npe;
    m[7] = km;
    for (int j = 0; j < nl; j++)
      if (slot_l[j] >= 0) {
        P.bnd_nodes.push_back(dev_l[j]);
        P.bnd_slot.push_back(slot_l[j]);
      }
    std::vector<int32_t> seen(nl, 0);
    for (int t = 0; t < nel; t++) {
      int e = order[es + t];
      P.elem_perm[es + t] = e;
      for (int j = 0; j < npe; j++) {
        const int l = local[conn[(size_t)npe * e + j]];
        const int k = seen[l]++;
        P.conn_local[(size_t)(es + t) * npe + j] = (uint16_t)l;
        P.csr_pos[(size_t)(es + t) * npe + j] = (uint16_t)(off[k] + l);
        P.corner[(size_t)(es + t) * npe + j] = (uint16_t)((l & 511) | ((k & 127) << 9));
      }
    }
    P.max_local_nodes = std::max(P.max_local_nodes, nl);
    P.max_block_elems = std::max(P.max_block_elems, nel);
    P.kmax = std::max(P.kmax, km);
    b_dev[b] = std::move(dev_l);
    b_slot[b] = std::move(slot_l);
    b_deg[b] = std::move(deg_l);
    b_off[b] = std::move(off);
  }
  // fixed-stride tables: local node l of block b at b*ml + l
  const int ml = P.max_local_nodes, ko = P.kmax + 1;
  P.nid_pad.assign((size_t)nb * ml, -1);
  P.slot_pad.assign((size_t)nb * ml, -1);
  P.deg_pad.assign((size_t)nb * ml, 0);
  P.off_pad.assign((size_t)nb * ko, 0);
  for (int b = 0; b < nb; b++) {
    for (size_t l = 0; l < b_dev[b].size(); l++) {
      P.nid_pad[(size_t)b * ml + l] = b_dev[b][l];
      P.slot_pad[(size_t)b * ml + l] = b_slot[b][l];
      P.deg_pad[(size_t)b * ml + l] = (uint16_t)b_deg[b][l];
    }
    for (int k = 0; k < ko; k++)
      P.off_pad[(size_t)b * ko + k] =
          (uint16_t)b_off[b][std::min<size_t>(k, b_off[b].size() - 1)];
  }
  return true;
}

}  // namespace cpi

using namespace cpi;

// ---------------------------------------------------------------------------
// C ABI: host-only utilities
// ---------------------------------------------------------------------------

extern "C" int cpi_greedy_coloring(int32_t ne, int32_t npe, const int32_t *conn,
                                   int32_t nn, int32_t *color) {
  if (ne < 0 || nn <= 0 || (npe != 4 && npe != 10) || !conn || !color) {
    set_global_error("cpi_greedy_coloring: invalid arguments");
    return CPI_ERR_INVALID;
  }
  for (long i = 0; i < (long)ne * npe; i++)
    if (conn[i] < 0 || conn[i] >= nn) {
      set_global_error("cpi_greedy_coloring: node index out of range");
      return CPI_ERR_MESH;
    }
  return greedy_coloring(ne, npe, conn, nn, color);
}

namespace {

// SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
// generators", OOPSLA 2014)
struct SplitMix64 {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// The 6 tets of a unit cell, one per ordering of the axes: the path
// (0,0,0) -> +e_a -> +e_b -> (1,1,1).  Corner offsets as bit masks x=1 y=2 z=4.
void kuhn_tets(int tets[6][4]) {
  static const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2},
                                  {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  for (int t = 0; t < 6; t++) {
    int v = 0;
    tets[t][0] = 0;
    for (int s = 0; s < 3; s++) {
      v |= 1 << perms[t][s];
      tets[t][s + 1] = v;
    }
  }
}

// corner pairs of the 6 mid-side nodes in C3D10 order
const int kEdge[6][2] = {{0, 1}, {1, 2}, {0, 2}, {0, 3}, {1, 3}, {2, 3}};

double det_corner(const double *x, const int32_t *c) {
  double e[3][3];
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < 3; k++) e[i][k] = x[3 * c[i + 1] + k] - x[3 * c[0] + k];
  return e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) -
         e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
         e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
}

// swap corners 0 and 1 (and the mid-sides that follow them)
void flip(int32_t *c, int npe) {
  std::swap(c[0], c[1]);
  if (npe == 10) {
    std::swap(c[5], c[6]);  // (1,2) <-> (0,2)
    std::swap(c[7], c[8]);  // (0,3) <-> (1,3)
  }
}

// Generic structured generator over a lattice of (n0+1)x(n1+1)x(n2+1)
// corner points (axis 0 optionally periodic).  pos(i,j,k) maps a corner to
// space; tet10 mid-sides are chord midpoints of the two corners of the edge.
template <class Pos>
int build_lattice(int n0, int n1, int n2, bool per0, bool quad, Pos pos,
                  int32_t *nn_out, int32_t *ne_out, double *x, int32_t *conn) {
  const int p0 = per0 ? n0 : n0 + 1;  // corner points along axis 0
  const int m0 = quad ? (per0 ? 2 * n0 : 2 * n0 + 1) : p0;
  const int m1 = quad ? 2 * n1 + 1 : n1 + 1;
  const int m2 = quad ? 2 * n2 + 1 : n2 + 1;
  const long nn = (long)m0 * m1 * m2;
  const long ne = 6L * n0 * n1 * n2;
  if (nn > INT32_MAX || ne > INT32_MAX) return CPI_ERR_INVALID;
  *nn_out = (int32_t)nn;
  *ne_out = (int32_t)ne;
  if (!x || !conn) return CPI_OK;
  const int s = quad ? 2 : 1;  // lattice stride of a corner
  auto id = [&](int a, int b, int c) -> int32_t {
    if (per0) a %= m0;
    return (int32_t)(a + (long)m0 * (b + (long)m1 * c));
  };
  // positions
  for (int c = 0; c < m2; c++)
    for (int b = 0; b < m1; b++)
      for (int a = 0; a < m0; a++) {
        double p[3];
        if (!quad || ((a % 2 == 0) && (b % 2 == 0) && (c % 2 == 0))) {
          pos(a / s, b / s, c / s, p);
        } else {
          double pl[3], ph[3];
          pos(a / 2, b / 2, c / 2, pl);
          pos((a + 1) / 2, (b + 1) / 2, (c + 1) / 2, ph);
          for (int k = 0; k < 3; k++) p[k] = 0.5 * (pl[k] + ph[k]);
        }
        double *dst = x + 3 * (long)id(a, b, c);
        dst[0] = p[0];
        dst[1] = p[1];
        dst[2] = p[2];
      }
  int tets[6][4];
  kuhn_tets(tets);
  const int npe = quad ? 10 : 4;
  long e = 0;
  for (int k = 0; k < n2; k++)
    for (int j = 0; j < n1; j++)
      for (int i = 0; i < n0; i++)
        for (int t = 0; t < 6; t++, e++) {
          int32_t *ce = conn + e * npe;
          int cc[4][3];
          for (int v = 0; v < 4; v++) {
            int m = tets[t][v];
            cc[v][0] = (i + (m & 1)) * s;
            cc[v][1] = (j + ((m >> 1) & 1)) * s;
            cc[v][2] = (k + ((m >> 2) & 1)) * s;
            ce[v] = id(cc[v][0], cc[v][1], cc[v][2]);
          }
          if (quad)
            for (int q = 0; q < 6; q++) {
              const int *A = cc[kEdge[q][0]], *B = cc[kEdge[q][1]];
              ce[4 + q] = id((A[0] + B[0]) / 2, (A[1] + B[1]) / 2, (A[2] + B[2]) / 2);
            }
          // orientation: the reference's linear basis needs
          // det[x1-x0, x2-x0, x3-x0] > 0; the quadratic corner map differs
          // from it by an even permutation, so the same test serves both.
          if (det_corner(x, ce) < 0.0) flip(ce, npe);
        }
  return CPI_OK;
}

}  // namespace

extern "C" int cpi_mesh_box(const cpi_box_spec *sp, int32_t *nn, int32_t *ne,
                            double *x, int32_t *conn) {
  if (!sp || !nn || !ne || sp->nx <= 0 || sp->ny <= 0 || sp->nz <= 0 ||
      !(sp->lx > 0) || !(sp->ly > 0) || !(sp->lz > 0) || sp->jitter < 0 ||
      sp->jitter >= 0.35) {
    set_global_error("cpi_mesh_box: invalid spec");
    return CPI_ERR_INVALID;
  }
  const bool quad = sp->quadratic != 0;
  const double h[3] = {sp->lx / sp->nx, sp->ly / sp->ny, sp->lz / sp->nz};
  const int n[3] = {sp->nx, sp->ny, sp->nz};
  // corner displacements drawn in corner-id order before placing nodes
  std::vector<double> jit;
  if (x && sp->jitter > 0) {
    const long nc = (long)(n[0] + 1) * (n[1] + 1) * (n[2] + 1);
    jit.assign(3 * nc, 0.0);
    SplitMix64 rng{sp->seed};
    for (int k = 0; k <= n[2]; k++)
      for (int j = 0; j <= n[1]; j++)
        for (int i = 0; i <= n[0]; i++) {
          bool interior = i > 0 && i < n[0] && j > 0 && j < n[1] && k > 0 && k < n[2];
          if (!interior) continue;
          long c = i + (long)(n[0] + 1) * (j + (long)(n[1] + 1) * k);
          for (int a = 0; a < 3; a++)
            jit[3 * c + a] = (2.0 * rng.uniform() - 1.0) * sp->jitter * h[a];
        }
  }
  auto pos = [&](int i, int j, int k, double *p) {
    const int ijk[3] = {i, j, k};
    long c = i + (long)(n[0] + 1) * (j + (long)(n[1] + 1) * k);
    for (int a = 0; a < 3; a++)
      p[a] = sp->origin[a] + ijk[a] * h[a] + (jit.empty() ? 0.0 : jit[3 * c + a]);
  };
  return build_lattice(sp->nx, sp->ny, sp->nz, false, quad, pos, nn, ne, x, conn);
}

extern "C" int cpi_mesh_cylinder(const cpi_cylinder_spec *sp, int32_t *nn,
                                 int32_t *ne, double *x, int32_t *conn) {
  if (!sp || !nn || !ne || sp->ntheta < 3 || sp->nz <= 0 || sp->nr <= 0 ||
      !(sp->radius > 0) || !(sp->length > 0) || !(sp->thickness > 0) ||
      sp->thickness >= 2 * sp->radius) {
    set_global_error("cpi_mesh_cylinder: invalid spec");
    return CPI_ERR_INVALID;
  }
  const double two_pi = 6.283185307179586476925286766559;
  auto pos = [&](int i, int j, int k, double *p) {
    double th = two_pi * (double)(i % sp->ntheta) / sp->ntheta;
    double r = sp->radius - 0.5 * sp->thickness + sp->thickness * k / sp->nr;
    p[0] = sp->origin[0] + r * std::cos(th);
    p[1] = sp->origin[1] + r * std::sin(th);
    p[2] = sp->origin[2] + sp->length * j / sp->nz;
  };
  return build_lattice(sp->ntheta, sp->nz, sp->nr, true, sp->quadratic != 0, pos,
                       nn, ne, x, conn);
}

// ---------------------------------------------------------------------------
// .inp reader / writer
// ---------------------------------------------------------------------------

namespace {

std::string trim(const std::string &s) {
  const char *ws = " \t\n\r\f\v";
  size_t a = s.find_first_not_of(ws);
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(ws);
  return s.substr(a, b - a + 1);
}

bool contains_nocase(const std::string &s, const char *pat) {
  std::string a = s, b = pat;
  for (auto &ch : a) ch = (char)std::tolower((unsigned char)ch);
  for (auto &ch : b) ch = (char)std::tolower((unsigned char)ch);
  return a.find(b) != std::string::npos;
}

struct InpData {
  std::vector<std::pair<int, std::array<double, 3>>> nodes;
  std::map<int, std::vector<int>> elems;
  std::map<std::string, std::vector<int>> sets;
};

}  // namespace


static int parse_inp(const char *path, InpData &d, std::string &err) {
  std::ifstream f(path);
  if (!f.is_open()) {
    err = std::string("Failed to open file: ") + path;
    return CPI_ERR_IO;
  }
  bool in_n = false, in_e = false, in_s = false;
  std::string set_name, line;
  long lineno = 0;
  try {
    while (std::getline(f, line)) {
      lineno++;
      line = trim(line);
      if (line.empty() || line[0] == '*') {
        in_n = line.find("*Node") != std::string::npos;
        in_e = line.find("*Element") != std::string::npos;
        in_s = line.find("*Nset") != std::string::npos;
        if (in_s) {
          size_t p = line.find("Nset=");
          if (p != std::string::npos) {
            set_name = line.substr(p + 5);
            d.sets[set_name] = {};
          }
        }
        continue;
      }
      if (in_n) {
        std::istringstream ss(line);
        std::string tok;
        std::getline(ss, tok, ',');
        int id = std::stoi(tok);
        std::array<double, 3> p;
        std::getline(ss, tok, ',');
        p[0] = std::stod(tok);
        std::getline(ss, tok, ',');
        p[1] = std::stod(tok);
        std::getline(ss, tok);
        p[2] = std::stod(tok);
        d.nodes.push_back({id, p});
      } else if (in_e) {
        std::istringstream ss(line);
        int id;
        if (!(ss >> id)) continue;  // unreadable element line: skipped
        std::string rest, tok;
        std::getline(ss, rest);
        std::istringstream rs(rest);
        std::vector<int> ids;
        while (std::getline(rs, tok, ',')) {
          if (tok.empty()) continue;
          std::istringstream ts(tok);
          int v;
          if (ts >> v) ids.push_back(v);
        }
        d.elems[id] = ids;
      } else if (in_s) {
        std::istringstream ss(line);
        std::string tok;
        while (std::getline(ss, tok, ',')) {
          tok = trim(tok);
          if (!tok.empty()) d.sets[set_name].push_back(std::stoi(tok));
        }
      }
    }
  } catch (const std::exception &ex) {
    err = std::string(path) + ":" + std::to_string(lineno) + ": cannot parse (" +
          ex.what() + ")";
    return CPI_ERR_IO;
  }
  return CPI_OK;
}

extern "C" int cpi_inp_read(const char *path, int32_t npe, cpi_inp_counts *cnt,
                            double *xloc, int32_t *conn, int32_t *set_starts,
                            int32_t *set_idx, int32_t *fixed, int32_t *slave,
                            char *names) {
  if (!path || !cnt || (npe != 4 && npe != 10)) {
    set_global_error("cpi_inp_read: invalid arguments");
    return CPI_ERR_INVALID;
  }
  InpData d;
  std::string err;
  int rc = parse_inp(path, d, err);
  if (rc) {
    set_global_error(err);
    return rc;
  }
  const int nn = (int)d.nodes.size(), ne = (int)d.elems.size();
  cnt->num_nodes = nn;
  cnt->num_elements = ne;
  cnt->num_node_sets = (int)d.sets.size();
  long tot = 0, nfix = 0, nsl = 0, nb = 0;
  for (auto &s : d.sets) {
    tot += (long)s.second.size();
    if (contains_nocase(s.first, "fixed")) nfix += (long)s.second.size();
    if (contains_nocase(s.first, "slave")) nsl += (long)s.second.size();
    nb += (long)s.first.size() + 1;
  }
  cnt->num_set_indices = (int32_t)tot;
  cnt->num_fixed_nodes = (int32_t)nfix;
  cnt->num_slave_nodes = (int32_t)nsl;
  cnt->names_bytes = (int32_t)nb;
  // ids are assumed 1-based and contiguous (mesh.h:199-212); check them so a
  // bad file is an error rather than an out-of-bounds write
  for (auto &n : d.nodes)
    if (n.first < 1 || n.first > nn) {
      set_global_error("cpi_inp_read: node ids are not 1.." + std::to_string(nn));
      return CPI_ERR_IO;
    }
  for (auto &e : d.elems) {
    if (e.first < 1 || e.first > ne || (int)e.second.size() < npe) {
      set_global_error("cpi_inp_read: element " + std::to_string(e.first) +
                       " has a bad id or fewer than " + std::to_string(npe) + " nodes");
      return CPI_ERR_IO;
    }
    for (int j = 0; j < npe; j++)
      if (e.second[j] < 1 || e.second[j] > nn) {
        set_global_error("cpi_inp_read: element " + std::to_string(e.first) +
                         " references a missing node");
        return CPI_ERR_IO;
      }
  }
  if (xloc)
    for (auto &n : d.nodes)
      for (int k = 0; k < 3; k++) xloc[3 * (long)(n.first - 1) + k] = n.second[k];
  if (conn)
    for (auto &e : d.elems)
      for (int j = 0; j < npe; j++) conn[(long)npe * (e.first - 1) + j] = e.second[j] - 1;
  long pos = 0, pf = 0, ps = 0, pn = 0;
  int si = 0;
  for (auto &s : d.sets) {
    if (set_starts) set_starts[si] = (int32_t)pos;
    bool isf = contains_nocase(s.first, "fixed"), iss = contains_nocase(s.first, "slave");
    for (int v : s.second) {
      if (set_idx) set_idx[pos] = v - 1;
      pos++;
      if (isf && fixed) fixed[pf] = v - 1;
      if (iss && slave) slave[ps] = v - 1;
      pf += isf;
      ps += iss;
    }
    if (names) {
      std::memcpy(names + pn, s.first.c_str(), s.first.size() + 1);
    }
    pn += (long)s.first.size() + 1;
    si++;
  }
  if (set_starts) set_starts[si] = (int32_t)pos;
  return CPI_OK;
}

extern "C" int cpi_inp_write(const char *path, const cpi_mesh_desc *m) {
  if (!path || !m || !m->xloc || !m->element_nodes ||
      (m->nodes_per_element != 4 && m->nodes_per_element != 10)) {
    set_global_error("cpi_inp_write: invalid arguments");
    return CPI_ERR_INVALID;
  }
  FILE *f = std::fopen(path, "w");
  if (!f) {
    set_global_error(std::string("cpi_inp_write: cannot open ") + path);
    return CPI_ERR_IO;
  }
  const int npe = m->nodes_per_element;
  std::fprintf(f, "*Heading\n generated by cppimpact_amd\n*Node\n");
  for (int i = 0; i < m->num_nodes; i++)
    std::fprintf(f, "%d, %.17g, %.17g, %.17g\n", i + 1, m->xloc[3 * i],
                 m->xloc[3 * i + 1], m->xloc[3 * i + 2]);
  std::fprintf(f, "*Element, Type=%s, Elset=Solid\n", npe == 4 ? "C3D4" : "C3D10");
  for (int e = 0; e < m->num_elements; e++) {
    std::fprintf(f, "%d", e + 1);
    for (int j = 0; j < npe; j++)
      std::fprintf(f, ", %d", m->element_nodes[(long)npe * e + j] + 1);
    std::fprintf(f, "\n");
  }
  auto nset = [&](const char *name, int n, const int32_t *ids) {
    if (n <= 0 || !ids) return;
    std::fprintf(f, "*Nset, Nset=%s\n", name);
    for (int i = 0; i < n; i++)
      std::fprintf(f, "%d%s", ids[i] + 1, (i % 16 == 15 || i + 1 == n) ? "\n" : ", ");
  };
  nset("fixed", m->num_fixed_nodes, m->fixed_nodes);
  nset("slave", m->num_slave_nodes, m->slave_nodes);
  int rc = std::fclose(f);
  if (rc != 0) {
    set_global_error(std::string("cpi_inp_write: write failed: ") + path);
    return CPI_ERR_IO;
  }
  return CPI_OK;
}
