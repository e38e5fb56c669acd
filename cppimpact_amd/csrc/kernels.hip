// HIP kernels of the explicit-dynamics step for gfx950 and their launchers.
//
// One loop iteration of Dynamics::solve (dynamics.h:471-511) is, with the
// LDS-assembled scatter (the default):
//   k_elem_block4 (tet4; with the surrogate the elastic pass, k_mlp_queue
//                 and k_elem_block4_walk),
//   k_elem_block10 (tet10, 4-point rules), k_elem_block (tet10, 5 points)
//                 one workgroup per element block: stage the block's nodes
//                 in LDS, compute element forces (and the stress update),
//                 sum them per node in a fixed order, finish the nodes owned
//                 by the block (wall, gravity, leapfrog) and write one
//                 partial force per shared node
//   k_node_bnd    one thread per shared node: sum its partials in block
//                 order, then wall, gravity and leapfrog
// With the colour scatter the element pass is one k_elem_direct launch per
// colour (plain read-modify-write, no two elements of a colour share a
// node); with the atomic scatter it is one launch using FP64 atomics; both
// are followed by k_node_all.
//
// Nodal state: 24-byte records pos = x and vv = v, and the lumped mass m
// (read only by the nodes' owners: a step never writes it).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "device.h"
#include "elem.h"
#include "kernels.h"
#include "mlp_infer.h"

namespace cpi {

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------

// Sum / min over the wave, the same bits on every lane and in every run:
// each row of 16 lanes is reduced with DPP moves, then the four row values
// (lanes 0, 16, 32, 48) are combined in a fixed order.  All 64 lanes must
// be active.
__device__ __forceinline__ double wave_sum(double v) {
  v = row_reduce(v, [](double a, double b) { return a + b; });
  return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}
__device__ __forceinline__ double wave_min(double v) {
  v = row_reduce(v, [](double a, double b) { return fmin(a, b); });
  return fmin(fmin(lane_value(v, 0), lane_value(v, 16)), fmin(lane_value(v, 32), lane_value(v, 48)));
}

// wave_sum of a term that is zero on every lane of most waves (plastic
// work, inverted elements, penalty energy, contact force, and the nodal
// terms in the waves of an element kernel that update no node): a uniform
// branch skips the reduction then
__device__ __forceinline__ double wave_sum_sparse(double v) {
  return __ballot(v != 0.0) ? wave_sum(v) : 0.0;
}

// Each wave owns one energy row: its lanes' accumulators are reduced with
// wave_sum / wave_min (fixed order, so deterministic) and lane 0 updates the
// row; no workgroup barrier is involved.  Cumulative terms (internal and
// plastic work, gravity work, inverted-element count) accumulate in the row
// every step; the instantaneous ones (kinetic and penalty energy, contact
// force, stable dt) are only reduced on the steps that end a batch, the only
// ones whose rows are summed.  The row of a wave is base + its wave index in
// the grid.  The cumulative terms of the row are read when the kernel starts
// (row_prefetch), so the flush at its end only writes.
struct RowPre {
  double wint, wpl, bad, wgrav;
};

__device__ __forceinline__ RowPre row_prefetch(const Params &p, int base) {
  // wave-uniform address: the values land in scalar registers
  const int row = __builtin_amdgcn_readfirstlane(base + (int)(threadIdx.x >> 6));
  const double *r = p.epart + (long)row * EP_N;
  return RowPre{r[EP_WINT], r[EP_WPL], r[EP_BAD], r[EP_WGRAV]};
}

// One wave's energy terms into its row: the cumulative ones added to what
// the row held (pre), and with ALL the instantaneous ones (kinetic,
// penalty, contact force, stable dt) stored.
template <bool ALL>
__device__ __forceinline__ void wave_flush(const Params &p, int base, const EAcc &ea,
                                           const RowPre &pre) {
  const double wint = wave_sum(ea.wint), wpl = wave_sum_sparse(ea.wpl),
               bad = wave_sum_sparse(ea.bad), wg = wave_sum_sparse(ea.wgrav);
  double ke = 0.0, pen = 0.0, fc = 0.0, md = 0.0;
  if (ALL) {
    ke = wave_sum_sparse(ea.ke);
    pen = wave_sum_sparse(ea.pen);
    fc = wave_sum_sparse(ea.fc);
    md = wave_min(ea.mindt);
  }
  if ((threadIdx.x & 63) == 0) {
    double *r = p.epart + (long)(base + (threadIdx.x >> 6)) * EP_N;
    r[EP_WINT] = pre.wint + wint;
    r[EP_WPL] = pre.wpl + wpl;
    r[EP_BAD] = pre.bad + bad;
    r[EP_WGRAV] = pre.wgrav + wg;
    if (ALL) {
      r[EP_MINDT] = md;
      r[EP_KE] = ke;
      r[EP_PEN] = pen;
      r[EP_FC] = fc;
    }
  }
}

// ---------------------------------------------------------------------------
// node update: a = -f/m + wall + g; x += du; v += dt a  (dynamics.h:497-503,
// wall.h:37-44, dynamics_kernels.h:150-168).  The wall sees x + du = x_{n+1}.
// A held node (fixed set, when enabled) has a = 0 and no contact; its
// velocity is zero from the start, so it stays where it is.
// ---------------------------------------------------------------------------
template <bool AUX>
__device__ __forceinline__ void node_update(const Params &p, int g, const double f[3],
                                            const double x[3], const double v[3], double m,
                                            double dt, EAcc &ea) {
  double du[3], a[3];
  const double minv = 1.0 / m;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    du[c] = dt * v[c];
    a[c] = (m != 0.0) ? minv * (-f[c]) : 0.0;
  }
  const bool held = p.fixed_mask && p.fixed_mask[g];
  if (p.wall_on && !held && (!p.contact_mask || p.contact_mask[g])) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      if (c == p.wall_dim) {
        double pos = x[c] + du[c];
        double d = (pos - p.wall_loc) * p.wall_normal;
        if (d < 0.0) {
          a[c] += -1 * minv * d * p.wall_nk;
          ea.pen += 0.5 * p.wall_k * d * d;
          ea.fc += -p.wall_k * d;
        }
      }
    }
  }
  double ke = 0.0, wg = 0.0, xn[3], vn[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    a[c] = held ? 0.0 : a[c] + p.grav[c];
    vn[c] = v[c] + dt * a[c];
    double vi = vn[c] - 0.5 * dt * a[c];
    ke += vi * vi;
    wg += p.grav[c] * du[c];
    xn[c] = x[c] + du[c];
  }
  p.pos[g] = Vec3{xn[0], xn[1], xn[2]};
  p.vv[g] = Vec3{vn[0], vn[1], vn[2]};
  if (AUX) {
    p.acc[g] = Node4{a[0], a[1], a[2], 0.0};
    p.du[g] = Node4{du[0], du[1], du[2], 0.0};
    p.xprev[g] = Node4{x[0], x[1], x[2], m};
  }
  ea.ke += 0.5 * m * ke;
  ea.wgrav += m * wg;
}

// A local node's force from its deg corner-force slots in the block's LDS
// (jagged-diagonal: the k-th at off[k] + j), summed in slot order in the
// element precision R: FP64 for FP64 elements, FP32 for FP32 ones, whose
// element forces carry FP32 rounding already (the nodal update that
// follows is FP64 either way).
template <typename R>
__device__ __forceinline__ void node_force_sum(const R *lf, const uint16_t *joff, int j, int deg,
                                               double f[3]) {
  R s[3] = {R(0), R(0), R(0)};
  for (int k = 0; k < deg; k++) {
    const int q = joff[k] + j;
    s[0] += lf[3 * q + 0];
    s[1] += lf[3 * q + 1];
    s[2] += lf[3 * q + 2];
  }
  f[0] = double(s[0]);
  f[1] = double(s[1]);
  f[2] = double(s[2]);
}

// One block's partial force on a shared node, stored in the element
// kernel's precision: FP64 records (24 B) for FP64 elements, FP32 ones
// (12 B) for FP32 elements, whose element forces carry FP32 rounding
// already.  The boundary-node kernel sums them in FP64.
template <typename R>
__device__ __forceinline__ void store_partial(const Params &p, int slot, const double f[3]) {
  if constexpr (std::is_same<R, float>::value)
    reinterpret_cast<Vec3f *>(p.partial)[slot] = Vec3f{float(f[0]), float(f[1]), float(f[2])};
  else
    p.partial[slot] = Vec3{f[0], f[1], f[2]};
}

// A workgroup's LDS node records, structure of arrays of ml doubles each:
// x, y, z, m, vx, vy, vz.  Lanes gathering random nodes read 8-byte words at
// an 8-byte stride, so the reads spread over the LDS banks (records of 64
// bytes put every node on one of four bank groups).
struct LNodes {
  double *d;
  int ml;
  __device__ __forceinline__ void stage(int j, const Vec3 &x, double m, const Vec3 &vv) const {
    d[j] = x.x;
    d[ml + j] = x.y;
    d[2 * ml + j] = x.z;
    d[3 * ml + j] = m;
    d[4 * ml + j] = vv.x;
    d[5 * ml + j] = vv.y;
    d[6 * ml + j] = vv.z;
  }
  __device__ __forceinline__ void get(int l, double x[3], double v[3]) const {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      x[c] = d[c * ml + l];
      v[c] = d[(4 + c) * ml + l];
    }
  }
  __device__ __forceinline__ double mass(int l) const { return d[3 * ml + l]; }
};

// bytes of the LDS regions of the workgroup element kernels: node records,
// corner forces (jagged-diagonal slots), the slot offsets off[k]
__host__ __device__ constexpr int lds_round16(long n) { return (int)(((n + 15) / 16) * 16); }
__host__ __device__ constexpr int lds_node_bytes(int ml) { return lds_round16(56L * ml); }
__host__ __device__ constexpr int lds_force_bytes(int rsz, int npe, int max_el) {
  return lds_round16((long)rsz * 3 * npe * max_el);
}
__host__ __device__ constexpr int lds_off_bytes(int kmax) { return lds_round16(2L * (kmax + 1)); }

// ---------------------------------------------------------------------------
// LDS-assembled element pass, tet10: loops over elements and nodes.  LDS:
// node records (LNodes), corner forces [max_el*NPE][3] R in jagged-diagonal
// slots (local node j's k-th contribution at off[k] + j), off[k].
// ---------------------------------------------------------------------------
template <int NPE, int RULE, typename R, int PHYS, bool AUX, int NT>
__global__ void __launch_bounds__(NT) k_elem_block(Params p) {
  extern __shared__ double smem[];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const RowPre row = row_prefetch(p, p.epart_base + b * (NT / 64));
  const int32_t *mt = p.meta + 8 * b;
  const int es = mt[0], nel = mt[1];
  const int ml = p.ml;
  const int32_t *nid = p.nid_pad + (long)b * ml;
  const LNodes ln{smem, ml};
  R *lf = (R *)((char *)smem + lds_node_bytes(ml));
  uint16_t *joff = (uint16_t *)((char *)lf + lds_force_bytes(sizeof(R), NPE, p.max_el));
  const double dt = p.ctl->dt;

  for (int j = threadIdx.x; j < ml; j += NT) {
    const int g = nid[j];
    if (g >= 0) ln.stage(j, p.pos[g], p.m[g], p.vv[g]);
  }
  for (int k = threadIdx.x; k <= p.kmax; k += NT) joff[k] = p.off_pad[(long)b * (p.kmax + 1) + k];
  __syncthreads();

  EAcc ea;
  for (int t = threadIdx.x; t < nel; t += NT) {
    const int e = es + t;
    const uint16_t *lc = p.conn_local + (long)e * NPE;
    const uint16_t *pos = p.csr_pos + (long)e * NPE;
    auto get = [&](int k, double xk[3], double vk[3]) { ln.get(lc[k], xk, vk); };
    // the element's corner slots belong to this thread alone: its forces
    // accumulate there point by point instead of in registers
#pragma unroll
    for (int k = 0; k < NPE; k++) {
      R *dst = lf + 3 * pos[k];
      dst[0] = dst[1] = dst[2] = R(0);
    }
    auto add = [&](int k, const R dn[3], const R s[6], R wd) {
      R *dst = lf + 3 * pos[k];
      R fk[3] = {dst[0], dst[1], dst[2]};
      add_bts<R>(fk, dn, s, wd);
      dst[0] = fk[0];
      dst[1] = fk[1];
      dst[2] = fk[2];
    };
    tet10_points<RULE, R, PHYS>(p, get, dt, e, add, ea);
  }
  __syncthreads();

  for (int j = threadIdx.x; j < ml; j += NT) {
    const int g = nid[j];
    if (g < 0) continue;
    const int deg = p.deg_pad[(long)b * ml + j];
    double f[3];
    node_force_sum<R>(lf, joff, j, deg, f);
    const int slot = p.slot_pad[(long)b * ml + j];
    if (slot < 0) {
      double x[3], v[3];
      ln.get(j, x, v);
      node_update<AUX>(p, g, f, x, v, ln.mass(j), dt, ea);
    } else {
      store_partial<R>(p, slot, f);
    }
  }
  wave_flush<AUX>(p, p.epart_base + b * (NT / 64), ea, row);
}

// tet10 on a 4-point rule: four lanes per element, lane t & 3 taking
// quadrature point t & 3, so a workgroup of NT threads holds at most NT/4
// elements, and the plan gives it at most NT local nodes, one per thread.
// As in the tet4 kernel every global load of the workgroup (the node's id,
// slot and contribution count, the node records and slot offsets, the
// element's ten corner words, the point's stress state with its ep only
// when the point's yielded bit says it has yielded) is issued before the
// first barrier; after it the element pass reads only LDS and the node pass
// only LDS and its registers.  Corner k's element force (summed over the
// element's four lanes in rule order) is stored by lane k & 3 into the
// corner's jagged-diagonal slot, off[rank] + local node; the nodes are then
// assembled as in k_elem_block, and the point's state is stored last.  LDS
// as k_elem_block.
template <int RULE, typename R, int PHYS, bool AUX, int NT, int WPE, bool STREAM>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE)))
k_elem_block10(Params p) {
  extern __shared__ double smem[];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int t = threadIdx.x, q = t & 3;
  const RowPre row = row_prefetch(p, p.epart_base + b * (NT / 64));
  const int32_t *mt = p.meta + 8 * b;
  const int es = mt[0], nel = mt[1];
  const LNodes ln{smem, p.ml};
  R *lf = (R *)((char *)smem + lds_node_bytes(p.ml));
  uint16_t *joff = (uint16_t *)((char *)lf + lds_force_bytes(sizeof(R), 10, p.max_el));
  const double dt = p.ctl->dt;
  // yielded bits, one per lane (point) of the wave: a point's ep is read
  // only when it has yielded before
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const bool has_word = PHYS == PHYS_EP && p.yielded && wv < p.ywords;
  const long yi = (long)b * p.ywords + wv;
  unsigned long long yw = ~0ull;
  if (has_word) yw = p.yielded[yi];
  // the loads in the order of the tet4 body (block4): node ids, then the
  // node records and slot offsets, then corner words and stress state, with
  // no branch between them (one offset per thread: kmax <= max_el <= NT/4)
  const long o = (long)b * p.ml + (t < p.ml ? t : p.ml - 1);
  const int slot_o = p.slot_pad[o], deg_o = p.deg_pad[o], g_o = p.nid_pad[o];
  const bool has_n = t < p.ml && g_o >= 0;
  const int g = has_n ? g_o : -1, slot = has_n ? slot_o : -1, deg = has_n ? deg_o : 0;
  const Vec3 xr = p.pos[has_n ? g_o : 0];
  const double mr = p.m[has_n && slot_o < 0 ? g_o : 0];
  const Vec3 vr = p.vv[has_n ? g_o : 0];
  const uint16_t jo = p.off_pad[(long)b * (p.kmax + 1) + (t < p.kmax ? t : p.kmax)];
  __builtin_amdgcn_sched_barrier(0);
  const bool has_e = (t >> 2) < nel;
  const int e = es + (t >> 2), ec = has_e ? e : es + nel - 1;
  // the element's ten corner words (local node | rank << 9), two per word
  // (the rows of 20 bytes are 4-byte aligned)
  uint32_t cw[5];
  {
    const uint32_t *cp = reinterpret_cast<const uint32_t *>(p.corner + 10L * ec);
#pragma unroll
    for (int k = 0; k < 5; k++) cw[k] = cp[k];
  }
  PreState<R> pre;
  if constexpr (PHYS != PHYS_COMPAT)
    load_state_sel<R, STREAM>(p, ec, q, (yw >> (t & 63)) & 1ull, pre);
  __builtin_amdgcn_sched_barrier(0);
  if (has_n) ln.stage(t, xr, mr, vr);
  if (t <= p.kmax) joff[t] = jo;
  __syncthreads();

  EAcc ea;
  // set where the point is computed: a lane without one never reads its
  // loaded state again, so nothing after the element pass waits for the
  // state loads (nor, by the in-order wait count, for the state stores)
  bool yielded = false;
  if (has_e) {
    auto half = [&](int k) { return (int)((cw[k >> 1] >> (16 * (k & 1))) & 0xffffu); };
    auto get = [&](int k, double xk[3], double vk[3]) { ln.get(half(k) & 511, xk, vk); };
    // lane q stores corners q, q + 4 and q + 8: their force slots
    int sl[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const uint32_t wj = (q >> 1) ? cw[j == 2 ? 4 : 2 * j + 1] : cw[2 * j];
      const int h = (int)((wj >> (16 * (q & 1))) & 0xffffu);
      sl[j] = (int)joff[h >> 9] + (h & 511);
    }
    auto put = [&](int k, const R f[3]) {
      if ((k & 3) == q) {
        R *dst = lf + 3 * sl[k >> 2];
        dst[0] = f[0];
        dst[1] = f[1];
        dst[2] = f[2];
      }
    };
    tet10_point_lane<RULE, R, PHYS, STREAM, !AUX>(p, get, dt, e, q, pre, put, ea);
    yielded = pre.ep > R(0);
  }
  if constexpr (PHYS == PHYS_EP) {
    const unsigned long long nb = __ballot(yielded);
    if (has_word && (t & 63) == 0 && nb != yw) p.yielded[yi] = nb;
  }
  __syncthreads();

  if (g >= 0) {
    double f[3];
    node_force_sum<R>(lf, joff, t, deg, f);
    if (slot < 0) {
      double x[3], v[3];
      ln.get(t, x, v);
      node_update<AUX>(p, g, f, x, v, ln.mass(t), dt, ea);
    } else {
      store_partial<R>(p, slot, f);
    }
  }
  wave_flush<AUX>(p, p.epart_base + b * (NT / 64), ea, row);
  // the point's state last (tet10_point_lane leaves it in pre), except in
  // the once-per-batch variant, whose outputs leave no registers for it
  if constexpr (PHYS != PHYS_COMPAT && !AUX)
    if (has_e) store_state<R, STREAM>(p, e, q, pre.s, pre.ep, pre.changed);
}

// tet4: the plan gives every workgroup at most NT elements and NT local
// nodes, so each thread owns one element and at most one node, and every
// global load of the workgroup (stress state, corner words, the
// contribution lists, node records, slots) is issued before the first
// barrier.  LDS: node records (LNodes), corner forces [max_el*4][3] R in
// jagged-diagonal slots, the slot offsets off[k].
//
// With the surrogate (PHYS_EP_MLP, FP32) a step is three launches:
//   elastic pass  this body over every block (WALK = false): a block whose
//                 trial stresses all stay inside the yield surface finishes
//                 here; a block with a yielding element writes nothing of
//                 its state, queues the network inputs of its yielding
//                 points (in thread order) and lists itself (one 64-bit
//                 atomic per deferred block: list slot and first queue row)
//   k_mlp_queue   the network on the whole queue (mlp_infer.h)
//   walk          this body (WALK = true) over the listed blocks: yielding
//                 elements take their queued outputs through mlp_apply
//                 (acceptance checks, FP64 exact fallback)
// A deferred block computes the same trial stresses, hence the same
// yielding set and queue order, in both launches, and which workgroup
// finishes a block never changes its results.
// CW: the element's corners as 16-bit words l | k << 9 (p.corner, 8 B per
// element), the slot off[k] + l taken from the LDS offsets after the
// barrier; otherwise the local ids and slots as two u16 arrays (16 B), for
// plans whose blocks exceed the words (a node in more than 128 elements of
// one block).
template <typename R, int PHYS, bool AUX, int NT, bool WALK, bool STREAM, bool CW>
__device__ __forceinline__ void block4(const Params &p, const int b, const unsigned qbase,
                                       int *wcnt, unsigned *misc) {
  extern __shared__ double smem[];
  const int t = threadIdx.x, w = t >> 6, lane = t & 63;
  const RowPre row = row_prefetch(p, p.epart_base + b * (NT / 64));
  const int32_t *mt = p.meta + 8 * b;
  const int es = mt[0], nel = mt[1];
  const LNodes ln{smem, p.ml};
  R *lf = (R *)((char *)smem + lds_node_bytes(p.ml));
  uint16_t *joff = (uint16_t *)((char *)lf + lds_force_bytes(sizeof(R), 4, p.max_el));
  const double dt = p.ctl->dt;

  // yielded bits of this wave's elements: a block has ywords words, one
  // per 64 elements of the largest block, and a wave past them (a block
  // plan of at most 320 elements per block) has no elements and no word
  const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
  const bool has_word = PHYS != PHYS_COMPAT && p.yielded && wv < p.ywords;
  const long yi = (long)b * p.ywords + wv;
  unsigned long long yw = ~0ull;
  if (has_word) yw = p.yielded[yi];
  // The loads of the workgroup, issued with no branch between them in the
  // order they are needed: the node ids, slots and degrees (fixed stride,
  // the record addresses depend on them), then the node records and the
  // slot offsets the first barrier waits for, then the corner words and
  // the stress state, which stay in flight across the barrier until the
  // element pass uses them.  Lanes past the block's nodes or elements load
  // the last valid entry and drop it.  One offset per thread covers the
  // block: kmax (the largest contribution count) <= max_el <= NT.
  const long o = (long)b * p.ml + (t < p.ml ? t : p.ml - 1);
  const int slot_o = p.slot_pad[o], deg_o = p.deg_pad[o], g_o = p.nid_pad[o];
  const bool has_n = t < p.ml && g_o >= 0;
  const int g = has_n ? g_o : -1, slot = has_n ? slot_o : -1, deg = has_n ? deg_o : 0;
  const Vec3 xr = p.pos[has_n ? g_o : 0];
  const double mr = p.m[has_n && slot_o < 0 ? g_o : 0];
  const Vec3 vr = p.vv[has_n ? g_o : 0];
  const uint16_t jo = p.off_pad[(long)b * (p.kmax + 1) + (t < p.kmax ? t : p.kmax)];
  __builtin_amdgcn_sched_barrier(0);
  const bool has_e = t < nel;
  const int e = es + t, ec = has_e ? e : es + nel - 1;
  const uint2 lc4 = *(const uint2 *)((CW ? p.corner : p.conn_local) + 4L * ec);
  uint2 ps4 = make_uint2(0u, 0u);
  if constexpr (!CW) ps4 = *(const uint2 *)(p.csr_pos + 4L * ec);
  PreState<R> pre;
  if constexpr (PHYS != PHYS_COMPAT)
    load_state_sel<R, STREAM>(p, ec, 0, (yw >> (t & 63)) & 1ull, pre);
  __builtin_amdgcn_sched_barrier(0);
  if (has_n) ln.stage(t, xr, mr, vr);
  if (t <= p.kmax) joff[t] = jo;
  if constexpr (WALK)
    if (t == 0) misc[0] = misc[1] = 0u;
  __syncthreads();

  EAcc ea;
  // corner k's 16-bit field of the two loaded words: its local id for the
  // gather and its slot for the force, decoded where they are used, or up
  // front in the walk (each form the one with fewer VGPRs)
  auto half = [&](const uint2 &u, int k) -> int {
    const unsigned v = k < 2 ? u.x : u.y;
    return (int)((k & 1) ? (v >> 16) : (v & 0xffffu));
  };
  auto lid = [&](int k) -> int { return CW ? (half(lc4, k) & 511) : half(lc4, k); };
  auto pos = [&](int k) -> int {
    const int h = half(lc4, k);
    return CW ? joff[h >> 9] + (h & 511) : half(ps4, k);
  };
  [[maybe_unused]] int l[4], ps[4];
  if constexpr (WALK) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      l[k] = lid(k);
      ps[k] = pos(k);
    }
  }
  auto get = [&](int k, double xk[3], double vk[3]) {
    if constexpr (WALK)
      ln.get(l[k], xk, vk);
    else
      ln.get(lid(k), xk, vk);
  };
  auto put = [&](int k, const R fk[3]) {
    R *dst;
    if constexpr (WALK)
      dst = lf + 3 * ps[k];
    else
      dst = lf + 3 * pos(k);
#pragma unroll
    for (int c = 0; c < 3; c++) dst[c] = fk[c];
  };
  [[maybe_unused]] R Ji[9], de[6], wd, snew[6], wk;
  [[maybe_unused]] bool yielding = false, ep_changed = false;
  // set where the element is computed (as in k_elem_block10)
  [[maybe_unused]] bool has_ep = false;
  if (has_e) {
    if constexpr (PHYS == PHYS_EP_MLP) {
      // an element whose trial stress stays inside the yield surface
      // finishes here; a yielding one waits for the network
      tet4_kin<R, PHYS_EP_MLP>(p, get, dt, Ji, de, wd, ea);
      wk = ep_trial<R>(p, de, pre.s, snew);
      if (!trial_yields<R>(p, snew, pre.ep)) {
#pragma unroll
        for (int c = 0; c < 6; c++) wk += snew[c] * de[c];
        ea.wint += double(wd * R(0.5) * wk);
        tet4_forces<R>(Ji, snew, wd, put);
      } else {
        yielding = true;
      }
    } else {
      tet4_core<R, PHYS, true, STREAM>(p, get, dt, e, put, ea, &pre);
      if constexpr (PHYS == PHYS_EP) has_ep = pre.ep > R(0);
    }
  }
  if constexpr (PHYS == PHYS_EP) {
    const unsigned long long nb = __ballot(has_ep);
    if (has_word && (t & 63) == 0 && nb != yw) p.yielded[yi] = nb;
  }
  if constexpr (PHYS == PHYS_EP_MLP) {
    // rank of each yielding element in thread order
    const uint64_t m = __ballot(yielding);
    if (lane == 0) wcnt[w] = __popcll(m);
    __syncthreads();
    int total = 0, off = 0;
#pragma unroll
    for (int i = 0; i < NT / 64; i++) {
      const int c = wcnt[i];
      off += i < w ? c : 0;
      total += c;
    }
    const int rank = off + __popcll(m & ((1ull << lane) - 1ull));
    if constexpr (!WALK) {
      if (total > 0) {
        // nothing of the block has reached global memory yet: list it and
        // queue its yielding points' network inputs
        if (t == 0) {
          const unsigned long long old =
              atomicAdd(p.mlp_ctr, (1ull << 32) | (unsigned long long)total);
          p.mlp_list[old >> 32] = make_int2(b, (int)(unsigned)old);
          misc[0] = (unsigned)old;
        }
        __syncthreads();
        if (yielding) {
          float x[kMlpIn];
          mlp_inputs<R>(*p.mlp, de, pre.s, pre.ep, x);
          mlp_store_inputs(p.mlp_rows + (size_t)(misc[0] + rank) * kMlpRowHalves, x);
        }
        return;
      }
      // stored here, not last: held through the node pass, the stress
      // would not fit the elastic pass's 64 VGPRs
      if (has_e) store_state<R, STREAM>(p, e, 0, snew, pre.ep, false);
    } else {
      int how = MLP_ELASTIC;
      if (yielding) {
        // the element again from the LDS node records (the same bits as
        // above; counted there), so nothing of it stays live across the
        // return map
        R Jy[9], dy[6], wy, sy[6];
        EAcc again;
        tet4_kin<R, PHYS_EP_MLP>(p, get, dt, Jy, dy, wy, again);
        R wky = ep_trial<R>(p, dy, pre.s, sy);
        R ep = pre.ep;
        const R q = yield_overshoot<R>(p, sy, ep);
        float y[kMlpOut];
        mlp_load_outputs(p.mlp_rows + (size_t)(qbase + rank) * kMlpRowHalves, y);
        const R dg = mlp_apply<R>(p, y, q, ep, sy, how);
#pragma unroll
        for (int c = 0; c < 6; c++) wky += sy[c] * dy[c];
        ea.wint += double(wy * R(0.5) * wky);
        if (dg > R(0))
          ea.wpl += double(wy * R(0.81649658092772603273) * (R(p.sy0) + R(p.H) * ep) * dg);
        tet4_forces<R>(Jy, sy, wy, put);
#pragma unroll
        for (int c = 0; c < 6; c++) snew[c] = sy[c];
        pre.ep = ep;
        ep_changed = true;
      }
      mlp_count(misc, how == MLP_ACCEPTED, how == MLP_FALLBACK);
      __syncthreads();
      if (t == 0 && p.mlp_stats && (misc[0] | misc[1])) {
        p.mlp_stats[2L * b] += misc[0];
        p.mlp_stats[2L * b + 1] += misc[1];
      }
    }
  } else {
    __syncthreads();
  }

  if (has_n) {
    double f[3];
    node_force_sum<R>(lf, joff, t, deg, f);
    if (slot < 0) {
      double x[3], v[3];
      ln.get(t, x, v);
      node_update<AUX>(p, g, f, x, v, ln.mass(t), dt, ea);
    } else {
      store_partial<R>(p, slot, f);
    }
  }
  wave_flush<AUX>(p, p.epart_base + b * (NT / 64), ea, row);
  // the walk stores the element's state last: registers that a store
  // reads must not be reused before it completes, so a store issued
  // earlier would hold up the node pass (the exact-return body stores it
  // in point_stress: held to the end, the FP64 kernel ran 1-2% slower)
  if constexpr (PHYS == PHYS_EP_MLP && WALK) {
    if (has_e) store_state<R, STREAM>(p, e, 0, snew, pre.ep, ep_changed);
  }
}

template <typename R, int PHYS, bool AUX, int NT, int WPE, bool STREAM, bool CW>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE)))
k_elem_block4(Params p) {
  __shared__ int wcnt[NT / 64];
  __shared__ unsigned misc[2];
  block4<R, PHYS, AUX, NT, false, STREAM, CW>(p, xcd_remap(blockIdx.x, gridDim.x), 0u, wcnt, misc);
}

// The surrogate step's third launch: one workgroup per block of the grid,
// workgroup i finishing the i-th deferred block listed this step, if there
// is one, so the hardware spreads a plastic zone over the CUs as it would
// the blocks of an element pass (the list order depends on atomic order,
// the results of a block do not).  A loop over several listed blocks per
// workgroup holds the kernel arguments in SGPRs across iterations and
// spills 40-50 VGPRs.
template <bool AUX, int NT, int WPE, bool STREAM, bool CW>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE)))
k_elem_block4_walk(Params p) {
  __shared__ int wcnt[NT / 64];
  __shared__ unsigned misc[2];
  if (blockIdx.x >= (unsigned)(*p.mlp_ctr >> 32)) return;
  const int2 ent = p.mlp_list[blockIdx.x];
  block4<float, PHYS_EP_MLP, AUX, NT, true, STREAM, CW>(p, ent.x, (unsigned)ent.y, wcnt, misc);
}

// The network on queue rows [0, n): n from the step's counter (ctr) or
// given.  Each workgroup with a tile to do copies the weights into LDS;
// its waves take tiles w, w + waves, ..., each reading the next tile's
// inputs before it runs the network on the current one, so the row loads
// (rows the previous launch wrote, from another XCD's L2 at best) overlap
// the MFMA chain.
template <int NT>
__global__ void __launch_bounds__(NT) k_mlp_queue(const MlpParams *P, uint16_t *rows,
                                                  const unsigned long long *ctr, int n_given) {
  __shared__ __attribute__((aligned(16))) MlpWeights W;
  constexpr int WPB = NT / 64;
  const unsigned n = ctr ? (unsigned)(*ctr & 0xffffffffull) : (unsigned)n_given;
  const int ntiles = (int)((n + 15u) >> 4);
  if ((int)blockIdx.x * WPB >= ntiles) return;
  mlp_lds_load(P->weights, &W);
  const int stride = (int)gridDim.x * WPB;
  int tile = (int)blockIdx.x * WPB + (int)(threadIdx.x >> 6);
  bf16x8 xb = mlp_tile_inputs(rows + (size_t)(tile < ntiles ? tile : 0) * 16 * kMlpRowHalves);
  __syncthreads();
  for (; tile < ntiles; tile += stride) {
    const int next = tile + stride < ntiles ? tile + stride : tile;
    const bf16x8 xn = mlp_tile_inputs(rows + (size_t)next * 16 * kMlpRowHalves);
    mlp_tile(W, xb, rows + (size_t)tile * 16 * kMlpRowHalves);
    xb = xn;
  }
}

template <bool AUX, typename R>
__global__ void __launch_bounds__(256) k_node_bnd(Params p) {
  using PR = typename std::conditional<std::is_same<R, float>::value, Vec3f, Vec3>::type;
  const PR *part = reinterpret_cast<const PR *>(p.partial);
  const int i = xcd_remap(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  const RowPre row = row_prefetch(p, p.epart_base + blockIdx.x * 4);
  const double dt = p.ctl->dt;
  EAcc ea;
  if (i < p.n_boundary) {
    const int g = p.n_interior + i;
    const Vec3 xr = p.pos[g], vv = p.vv[g];
    const double m = p.m[g];
    // partials in slot order, loaded four at a time
    double f[3] = {0.0, 0.0, 0.0};
    const int s1 = p.slot_ptr[i + 1];
    int s = p.slot_ptr[i];
    auto add = [&](const PR &q) {
      f[0] += double(q.x);
      f[1] += double(q.y);
      f[2] += double(q.z);
    };
    for (; s + 4 <= s1; s += 4) {
      const PR q0 = part[s], q1 = part[s + 1], q2 = part[s + 2], q3 = part[s + 3];
      add(q0);
      add(q1);
      add(q2);
      add(q3);
    }
    if (s + 2 <= s1) {
      const PR q0 = part[s], q1 = part[s + 1];
      add(q0);
      add(q1);
      s += 2;
    }
    if (s < s1) add(part[s]);
    const double x[3] = {xr.x, xr.y, xr.z}, v[3] = {vv.x, vv.y, vv.z};
    node_update<AUX>(p, g, f, x, v, m, dt, ea);
  }
  wave_flush<AUX>(p, p.epart_base + blockIdx.x * 4, ea, row);
  // the last launch of the step: the surrogate's queue and block list start
  // empty in the next one
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    p.ctl->step += 1;
    if (p.mlp_ctr) *p.mlp_ctr = 0ull;
  }
}

// ---------------------------------------------------------------------------
// colour / atomic element pass: one thread per element, nodes from HBM
// ---------------------------------------------------------------------------
template <int NPE, int RULE, typename R, int PHYS, bool ATOMIC, int NT>
__global__ void __launch_bounds__(NT) k_elem_direct(Params p) {
  const int t = blockIdx.x * NT + threadIdx.x;
  const RowPre row = row_prefetch(p, p.epart_base + blockIdx.x * (NT / 64));
  const double dt = p.ctl->dt;
  EAcc ea;
  if (t < p.elem_count) {
    const int e = p.elem_base + t;
    int nd[NPE];
#pragma unroll
    for (int k = 0; k < NPE; k++) nd[k] = p.conn[(long)k * p.ne + e];
    auto get = [&](int k, double xk[3], double vk[3]) {
      const Vec3 a = p.pos[nd[k]], b = p.vv[nd[k]];
      xk[0] = a.x;
      xk[1] = a.y;
      xk[2] = a.z;
      vk[0] = b.x;
      vk[1] = b.y;
      vk[2] = b.z;
    };
    auto put = [&](int k, const R fk[3]) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        if (ATOMIC)
          unsafeAtomicAdd(&p.f[c][nd[k]], double(fk[c]));
        else
          p.f[c][nd[k]] += double(fk[c]);
      }
    };
    elem_core<NPE, RULE, R, PHYS>(p, get, dt, e, put, ea);
  }
  wave_flush<true>(p, p.epart_base + blockIdx.x * (NT / 64), ea, row);
}

template <bool AUX>
__global__ void __launch_bounds__(256) k_node_all(Params p) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  const RowPre row = row_prefetch(p, p.epart_base + blockIdx.x * 4);
  const double dt = p.ctl->dt;
  EAcc ea;
  if (g < p.nn) {
    double f[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      f[c] = p.f[c][g];
      p.f[c][g] = 0.0;
    }
    const Vec3 xr = p.pos[g], vv = p.vv[g];
    const double x[3] = {xr.x, xr.y, xr.z}, v[3] = {vv.x, vv.y, vv.z};
    node_update<AUX>(p, g, f, x, v, p.m[g], dt, ea);
  }
  wave_flush<AUX>(p, p.epart_base + blockIdx.x * 4, ea, row);
  if (blockIdx.x == 0 && threadIdx.x == 0) p.ctl->step += 1;
}

// initial update with du = 0 (f_int = 0) and the half-kick
// (dynamics.h:454-463): a0 = wall(X0) + g; v += dt/2 a0
__global__ void __launch_bounds__(256) k_node_init(Params p) {
  const int g = blockIdx.x * 256 + threadIdx.x;
  const RowPre row = row_prefetch(p, blockIdx.x * 4);
  const double dt = p.ctl->dt;
  EAcc ea;
  if (g < p.nn) {
    const Vec3 xr = p.pos[g], vv = p.vv[g];
    const double m = p.m[g];
    const bool held = p.fixed_mask && p.fixed_mask[g];
    const double x[3] = {xr.x, xr.y, xr.z};
    const double v[3] = {held ? 0.0 : vv.x, held ? 0.0 : vv.y, held ? 0.0 : vv.z};
    double a[3] = {0.0, 0.0, 0.0}, ke = 0.0, vn[3];
    if (p.wall_on && !held && (!p.contact_mask || p.contact_mask[g])) {
#pragma unroll
      for (int c = 0; c < 3; c++)
        if (c == p.wall_dim) {
          double d = (x[c] - p.wall_loc) * p.wall_normal;
          if (d < 0.0) {
            a[c] += -1 * (1 / m) * d * p.wall_nk;
            ea.pen += 0.5 * p.wall_k * d * d;
            ea.fc += -p.wall_k * d;
          }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
      a[c] = held ? 0.0 : a[c] + p.grav[c];
      ke += v[c] * v[c];
      vn[c] = v[c] + 0.5 * dt * a[c];
    }
    p.vv[g] = Vec3{vn[0], vn[1], vn[2]};
    p.acc[g] = Node4{a[0], a[1], a[2], 0.0};
    p.du[g] = Node4{0.0, 0.0, 0.0, 0.0};
    p.xprev[g] = Node4{xr.x, xr.y, xr.z, m};
    ea.ke = 0.5 * m * ke;
  }
  wave_flush<true>(p, blockIdx.x * 4, ea, row);
}

// Energy rows are summed in two fixed-order stages: kEredBlocks workgroups
// each reduce a contiguous chunk of rows into p.ered, then one wave sums
// those and writes the energy row of this step.
constexpr int kEredBlocks = 64;

__global__ void __launch_bounds__(256) k_energy_reduce1(Params p) {
  __shared__ double red[256 * EP_N];
  const int chunk = (p.n_epart + kEredBlocks - 1) / kEredBlocks;
  const int r0 = blockIdx.x * chunk, r1 = min(p.n_epart, r0 + chunk);
  double v[EP_N] = {0, 0, 1e300, 0, 0, 0, 0, 0};
  for (int r = r0 + threadIdx.x; r < r1; r += 256) {
    const double *src = p.epart + (long)r * EP_N;
    // a stable-dt entry of 0 is a row no kernel has written since the rows
    // were cleared (cpi_init's own reduction, rows only the half-kick
    // writes); it takes no part in the minimum.  A degenerate element
    // (limit 0) is caught by the inverted-element count instead.
#pragma unroll
    for (int k = 0; k < EP_N; k++)
      v[k] = (k == EP_MINDT) ? fmin(v[k], src[k] > 0.0 ? src[k] : 1e300) : v[k] + src[k];
  }
#pragma unroll
  for (int k = 0; k < EP_N; k++) red[threadIdx.x * EP_N + k] = v[k];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int k = 0; k < EP_N; k++) {
        double o = red[(threadIdx.x + h) * EP_N + k];
        double &d = red[threadIdx.x * EP_N + k];
        d = (k == EP_MINDT) ? fmin(d, o) : d + o;
      }
    __syncthreads();
  }
  if (threadIdx.x < EP_N) p.ered[blockIdx.x * EP_N + threadIdx.x] = red[threadIdx.x];
}

__global__ void __launch_bounds__(64) k_energy_reduce2(Params p) {
  if (threadIdx.x == 0) {
    double s[EP_N] = {0, 0, 1e300, 0, 0, 0, 0, 0};
    for (int b = 0; b < kEredBlocks; b++)
      for (int k = 0; k < EP_N; k++) {
        const double o = p.ered[b * EP_N + k];
        s[k] = (k == EP_MINDT) ? fmin(s[k], o) : s[k] + o;
      }
    Ctl *c = p.ctl;
    double row[8];
    row[0] = s[EP_KE];
    row[1] = s[EP_WINT];
    row[2] = s[EP_WPL];
    row[3] = s[EP_PEN];
    row[4] = s[EP_WGRAV];
    row[5] = s[EP_FC];
    if (c->step == 0) c->e0 = row[0] + row[3];
    row[6] = row[0] + row[1] + row[3] - row[4] - c->e0;
    row[7] = s[EP_MINDT];
    bool finite = true;
    for (int k = 0; k < 7; k++) {
      c->row[k] = row[k];
      finite = finite && isfinite(row[k]);
    }
    c->row[7] = row[7];
    if (!finite) c->flags |= 1u;
    if (s[EP_BAD] > 0.0) c->flags |= 2u;
    // the elastoplastic kernels reduce the element CFL limit of the
    // configuration each step started from (the others leave 1e300): a
    // fixed dt above it, or above the wall spring's limit, is unstable
    if (row[7] < 1e300 && !(c->dt <= fmin(row[7], c->dt_contact))) c->flags |= 4u;
    if (p.hist && p.hist_cap > 0) {
      double *h = p.hist + (c->nrows % p.hist_cap) * 8;
      for (int k = 0; k < 8; k++) h[k] = row[k];
      c->nrows += 1;
    }
  }
}

// ---------------------------------------------------------------------------
// lumped mass, into m
// ---------------------------------------------------------------------------
template <int NPE, int RULE, int NT>
__global__ void __launch_bounds__(NT) k_mass_block(Params p, int kind, double rho) {
  extern __shared__ double smem[];
  const int b = blockIdx.x;
  const int32_t *mt = p.meta + 8 * b;
  const int es = mt[0], nel = mt[1];
  const int ml = p.ml;
  const int32_t *nid = p.nid_pad + (long)b * ml;
  double *lx = smem;               // x, y, z: ml each
  double *lm = smem + 3 * ml;      // corner masses, jagged-diagonal slots
  uint16_t *joff = (uint16_t *)(lm + (long)p.max_el * NPE);
  for (int j = threadIdx.x; j < ml; j += NT) {
    const int g = nid[j];
    if (g >= 0) {
      const Vec3 q = p.pos[g];
      lx[j] = q.x;
      lx[ml + j] = q.y;
      lx[2 * ml + j] = q.z;
    }
  }
  for (int k = threadIdx.x; k <= p.kmax; k += NT) joff[k] = p.off_pad[(long)b * (p.kmax + 1) + k];
  __syncthreads();
  for (int t = threadIdx.x; t < nel; t += NT) {
    const int e = es + t;
    double xs[NPE][3], me[NPE];
#pragma unroll
    for (int k = 0; k < NPE; k++) {
      const int l = p.conn_local[(long)e * NPE + k];
#pragma unroll
      for (int c = 0; c < 3; c++) xs[k][c] = lx[c * ml + l];
    }
    elem_mass<NPE, RULE>(xs, kind, rho, me);
#pragma unroll
    for (int k = 0; k < NPE; k++) lm[p.csr_pos[(long)e * NPE + k]] = me[k];
  }
  __syncthreads();
  for (int j = threadIdx.x; j < ml; j += NT) {
    const int g = nid[j];
    if (g < 0) continue;
    double sum = 0.0;
    const int deg = p.deg_pad[(long)b * ml + j];
    for (int k = 0; k < deg; k++) sum += lm[joff[k] + j];
    const int slot = p.slot_pad[(long)b * ml + j];
    if (slot < 0)
      p.m[g] = sum;
    else
      p.partial[slot].x = sum;
  }
}

__global__ void __launch_bounds__(256) k_mass_bnd(Params p) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= p.n_boundary) return;
  double s = 0.0;
  for (int q = p.slot_ptr[i]; q < p.slot_ptr[i + 1]; q++) s += p.partial[q].x;
  p.m[p.n_interior + i] = s;
}

template <int NPE, int RULE, bool ATOMIC>
__global__ void __launch_bounds__(256) k_mass_direct(Params p, int kind, double rho) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= p.elem_count) return;
  const int e = p.elem_base + t;
  int nd[NPE];
  double xs[NPE][3], me[NPE];
#pragma unroll
  for (int k = 0; k < NPE; k++) {
    nd[k] = p.conn[(long)k * p.ne + e];
    const Vec3 q = p.pos[nd[k]];
    xs[k][0] = q.x;
    xs[k][1] = q.y;
    xs[k][2] = q.z;
  }
  elem_mass<NPE, RULE>(xs, kind, rho, me);
#pragma unroll
  for (int k = 0; k < NPE; k++) {
    if (ATOMIC)
      unsafeAtomicAdd(&p.m[nd[k]], me[k]);
    else
      p.m[nd[k]] += me[k];
  }
}

// ---------------------------------------------------------------------------
// stable dt of the current configuration, one minimum per workgroup
// ---------------------------------------------------------------------------
template <int NPE>
__global__ void __launch_bounds__(256) k_elem_dt(Params p, double *out) {
  __shared__ double red[4];
  const int e = blockIdx.x * 256 + threadIdx.x;
  double d = 1e300;
  if (e < p.ne) {
    double xs[4][3];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const Vec3 q = p.pos[p.conn[(long)k * p.ne + e]];
      xs[k][0] = q.x;
      xs[k][1] = q.y;
      xs[k][2] = q.z;
    }
    // linear corner order: tet4 (0,1,2,3); tet10 corners (2,0,1,3)
    const int o = NPE == 4 ? 0 : 2, i1 = NPE == 4 ? 1 : 0, i2 = NPE == 4 ? 2 : 1;
    double e1[3], e2[3], e3[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      e1[c] = xs[i1][c] - xs[o][c];
      e2[c] = xs[i2][c] - xs[o][c];
      e3[c] = xs[3][c] - xs[o][c];
    }
    double v6 = e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
    d = corner_dt<double>(e1, e2, e3, v6, p.wave_speed,
                          NPE == 4 ? kTet4LengthFactor : kTet10LengthFactor);
  }
  d = wave_min(d);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = d;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
}

// ---------------------------------------------------------------------------
// per-element strain/stress at xi = 0 from the total displacement
// x_n + du - X0 with B(x_n) (dynamics_kernels.h:117-130, analysis.h:479-503)
// ---------------------------------------------------------------------------
template <int NPE>
__global__ void __launch_bounds__(256) k_stress_elem(Params p, double *out) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= p.ne) return;
  double xs[NPE][3], u[NPE][3];
#pragma unroll
  for (int k = 0; k < NPE; k++) {
    const int n = p.conn[(long)k * p.ne + e];
    const Node4 xp = p.xprev[n], d = p.du[n], x0 = p.X0[n];
    xs[k][0] = xp.x;
    xs[k][1] = xp.y;
    xs[k][2] = xp.z;
    u[k][0] = d.x + xp.x - x0.x;
    u[k][1] = d.y + xp.y - x0.y;
    u[k][2] = d.z + xp.z - x0.z;
  }
  double G[NPE * 3];
  if constexpr (NPE == 4) {
    const double g4[12] = {-1, -1, -1, 1, 0, 0, 0, 1, 0, 0, 0, 1};
#pragma unroll
    for (int i = 0; i < 12; i++) G[i] = g4[i];
  } else {
    tet10_grad(0.0, 0.0, 0.0, G);
  }
  double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < NPE; i++)
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int c = 0; c < 3; c++) J[a * 3 + c] += G[3 * i + c] * xs[i][a];
  double Ji[9];
  inv3<double>(J, Ji);
  double de[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
  for (int i = 0; i < NPE; i++) {
    double dn[3];
#pragma unroll
    for (int c = 0; c < 3; c++)
      dn[c] = Ji[c] * G[3 * i] + Ji[3 + c] * G[3 * i + 1] + Ji[6 + c] * G[3 * i + 2];
    de[0] += dn[0] * u[i][0];
    de[1] += dn[1] * u[i][1];
    de[2] += dn[2] * u[i][2];
    de[3] += dn[1] * u[i][0] + dn[0] * u[i][1];
    de[4] += dn[2] * u[i][1] + dn[1] * u[i][2];
    de[5] += dn[2] * u[i][0] + dn[0] * u[i][2];
  }
  double *o = out + 12L * e;
#pragma unroll
  for (int c = 0; c < 6; c++) o[c] = de[c];
  o[6] = p.c11 * de[0] + p.c12 * (de[1] + de[2]);
  o[7] = p.c11 * de[1] + p.c12 * (de[0] + de[2]);
  o[8] = p.c11 * de[2] + p.c12 * (de[0] + de[1]);
  o[9] = p.c44 * de[3];
  o[10] = p.c44 * de[4];
  o[11] = p.c44 * de[5];
}

// ---------------------------------------------------------------------------
// known-answer strain energy and volume on the reference positions
// (dynamics.h:303-412, analysis.h:384-477): W_e = 1/2 sum_q w detJ de.D.de,
// V_e = sum_q w detJ
// ---------------------------------------------------------------------------
template <int NPE, int RULE>
__global__ void __launch_bounds__(256) k_debug_strain(Params p, double alpha, int dcase,
                                                       double nu, double *out) {
  __shared__ double red[8];
  const int e = blockIdx.x * 256 + threadIdx.x;
  double W = 0.0, V = 0.0;
  if (e < p.ne) {
    double xs[NPE][3], u[NPE][3];
#pragma unroll
    for (int k = 0; k < NPE; k++) {
      const Node4 q = p.X0[p.conn[(long)k * p.ne + e]];
      xs[k][0] = q.x;
      xs[k][1] = q.y;
      xs[k][2] = q.z;
      const double x = q.x;
      const double s = dcase == 0 ? alpha * x : alpha * 0.5 * x * x;
      u[k][0] = s;
      u[k][1] = -s * nu;
      u[k][2] = -s * nu;
    }
#pragma unroll
    for (int q = 0; q < Quad<RULE>::n; q++) {
      double G[NPE * 3];
      if constexpr (NPE == 4) {
        const double g4[12] = {-1, -1, -1, 1, 0, 0, 0, 1, 0, 0, 0, 1};
#pragma unroll
        for (int i = 0; i < 12; i++) G[i] = g4[i];
      } else {
        tet10_grad(Quad<RULE>::pt(q, 0), Quad<RULE>::pt(q, 1), Quad<RULE>::pt(q, 2), G);
      }
      double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int i = 0; i < NPE; i++)
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
          for (int c = 0; c < 3; c++) J[a * 3 + c] += G[3 * i + c] * xs[i][a];
      double Ji[9];
      double detJ = inv3<double>(J, Ji);
      double de[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
      for (int i = 0; i < NPE; i++) {
        double dn[3];
#pragma unroll
        for (int c = 0; c < 3; c++)
          dn[c] = Ji[c] * G[3 * i] + Ji[3 + c] * G[3 * i + 1] + Ji[6 + c] * G[3 * i + 2];
        de[0] += dn[0] * u[i][0];
        de[1] += dn[1] * u[i][1];
        de[2] += dn[2] * u[i][2];
        de[3] += dn[1] * u[i][0] + dn[0] * u[i][1];
        de[4] += dn[2] * u[i][1] + dn[1] * u[i][2];
        de[5] += dn[2] * u[i][0] + dn[0] * u[i][2];
      }
      double s0 = p.c11 * de[0] + p.c12 * (de[1] + de[2]);
      double s1 = p.c11 * de[1] + p.c12 * (de[0] + de[2]);
      double s2 = p.c11 * de[2] + p.c12 * (de[0] + de[1]);
      double w = Quad<RULE>::w(q) * detJ;
      W += 0.5 * w *
           (s0 * de[0] + s1 * de[1] + s2 * de[2] +
            p.c44 * (de[3] * de[3] + de[4] * de[4] + de[5] * de[5]));
      V += w;
    }
  }
  W = wave_sum(W);
  V = wave_sum(V);
  if ((threadIdx.x & 63) == 0) {
    red[2 * (threadIdx.x >> 6)] = W;
    red[2 * (threadIdx.x >> 6) + 1] = V;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = red[0] + red[2] + red[4] + red[6];
    out[2 * blockIdx.x + 1] = red[1] + red[3] + red[5] + red[7];
  }
}

// ---------------------------------------------------------------------------
// streaming copy (bandwidth probe and counter calibration)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_stream_copy(const Node4 *__restrict__ src,
                                                     Node4 *__restrict__ dst, long n) {
  const long stride = (long)gridDim.x * 256;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------

// Events the next element or node kernel launch records at its own start
// and end (hipExtLaunchKernel), for cpi_profile_kernels: the kernel's
// execution time without the event packets' own latency.  Null: a plain
// launch.  Per host thread.
static thread_local hipEvent_t t_ev_start = nullptr, t_ev_stop = nullptr;

void set_launch_events(hipEvent_t start, hipEvent_t stop) {
  t_ev_start = start;
  t_ev_stop = stop;
}

template <typename K, typename... A>
static inline void timed_launch(K kern, dim3 grid, dim3 block, size_t smem, hipStream_t s,
                                A... args) {
  if (t_ev_start || t_ev_stop)
    hipExtLaunchKernelGGL(kern, grid, block, (std::uint32_t)smem, s, t_ev_start, t_ev_stop, 0,
                          args...);
  else
    hipLaunchKernelGGL(kern, grid, block, smem, s, args...);
}

#define CPI_LAUNCH(kern, grid, block, smem, stream, ...)                      \
  do {                                                                        \
    timed_launch((kern), dim3(grid), dim3(block), (smem), (stream), __VA_ARGS__); \
    return hipGetLastError();                                                 \
  } while (0)

// threads per workgroup of the block kernels: tet4 one per element (384),
// tet10 on a 4-point rule one per point (512 for 128 elements), tet10 on
// the 5-point rule one per element (128)
template <int NPE, int RULE> constexpr int block_threads() {
  return NPE == 4 ? 384 : (Quad<RULE>::n == 4 ? 512 : 128);
}
// default elements per workgroup
constexpr int block_elements(int npe) { return npe == 4 ? 384 : 128; }
// minimum waves per SIMD of the per-point tet10 kernel
template <typename R> constexpr int tet10_waves() { return 4; }

// minimum waves per SIMD the tet4 element kernels are compiled for: FP64
// fits 96 VGPRs (three 6-wave workgroups per CU), FP32 64 (five per CU);
// the surrogate step's elastic pass has the plain FP32 kernel's structure
// and budget.  Its walk (network outputs, acceptance checks, the FP64
// fallback) spills 19 VGPRs at 64 and none at 80: six waves per SIMD, four
// workgroups per CU.
template <typename R, int PHYS> constexpr int tet4_waves() {
  return std::is_same<R, double>::value ? 5 : 8;
}
constexpr int kMlpWalkWaves = 6;

// workgroups of k_mlp_queue in the step (four waves each, every wave
// looping over tiles): eight per CU, eight waves per SIMD at 32 VGPRs and
// 15 KB of LDS each
constexpr int kMlpQueueGrid = 2048;
constexpr int kMlpQueueThreads = 256;

// LDS of the workgroup element kernels: node records, corner forces, the
// slot offsets
size_t elem_block_smem(int npe, int prec, int, int ml, int max_el, int kmax) {
  const int rsz = prec == 1 ? 4 : 8;
  return (size_t)lds_node_bytes(ml) + lds_force_bytes(rsz, npe, max_el) + lds_off_bytes(kmax);
}

int elem_block_threads(int npe, int rule) {
  if (npe == 4) return block_threads<4, RULE_CENTROID>();
  return rule == RULE_REF5 ? block_threads<10, RULE_REF5>() : block_threads<10, RULE_EXACT4>();
}

int elem_block_default_elements(int npe) { return block_elements(npe); }

// dynamic LDS above 64 KiB needs an explicit opt-in per kernel
template <typename K>
static hipError_t lds_opt_in(K kernel, size_t smem) {
  if (smem <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
}

// the tet4 kernels' LDS opt-in, cached or streamed state, corner words or
// the two u16 arrays
template <typename R, int PHYS, bool STREAM, bool CW>
static hipError_t tet4_opt_in(size_t smem) {
  constexpr int NT = block_threads<4, RULE_CENTROID>();
  constexpr int W = tet4_waves<R, PHYS>();
  hipError_t e = lds_opt_in(k_elem_block4<R, PHYS, true, NT, W, STREAM, CW>, smem);
  if (e == hipSuccess) e = lds_opt_in(k_elem_block4<R, PHYS, false, NT, W, STREAM, CW>, smem);
  if constexpr (PHYS == PHYS_EP_MLP) {
    if (e == hipSuccess)
      e = lds_opt_in(k_elem_block4_walk<true, NT, kMlpWalkWaves, STREAM, CW>, smem);
    if (e == hipSuccess)
      e = lds_opt_in(k_elem_block4_walk<false, NT, kMlpWalkWaves, STREAM, CW>, smem);
  }
  return e;
}

// the LDS opt-in (prepare_elem_block) is done once at set-up, outside any
// stream capture
// the tet4 launches, with the stress state cached or streamed (STREAM),
// the corners as 16-bit words or two u16 arrays (CW)
template <typename R, int PHYS, bool STREAM, bool CW>
static hipError_t elem_block4_t(const Params &p, bool aux, int nblocks, size_t smem,
                                hipStream_t s, int part) {
  constexpr int NT = block_threads<4, RULE_CENTROID>();
  constexpr int W = tet4_waves<R, PHYS>();
  if constexpr (PHYS == PHYS_EP_MLP) {
    // the elastic pass, the network on the queue, the walk over the
    // deferred blocks (part 1, 2, 3 alone for per-launch timing)
    if (part == 0 || part == 1) {
      if (aux)
        timed_launch((k_elem_block4<R, PHYS, true, NT, W, STREAM, CW>), dim3(nblocks), dim3(NT), smem,
                     s, p);
      else
        timed_launch((k_elem_block4<R, PHYS, false, NT, W, STREAM, CW>), dim3(nblocks), dim3(NT),
                     smem, s, p);
    }
    if (part == 0 || part == 2) {
      const int tiles = (p.ne + 15) / 16;
      const int grid = std::max(1, std::min(kMlpQueueGrid, (tiles + 3) / 4));
      timed_launch((k_mlp_queue<kMlpQueueThreads>), dim3(grid), dim3(kMlpQueueThreads), 0, s,
                   p.mlp, p.mlp_rows, (const unsigned long long *)p.mlp_ctr, 0);
    }
    if (part == 0 || part == 3) {
      const int grid = nblocks;
      if (aux)
        timed_launch((k_elem_block4_walk<true, NT, kMlpWalkWaves, STREAM, CW>), dim3(grid), dim3(NT),
                     smem, s, p);
      else
        timed_launch((k_elem_block4_walk<false, NT, kMlpWalkWaves, STREAM, CW>), dim3(grid), dim3(NT),
                     smem, s, p);
    }
    return hipGetLastError();
  } else {
    if (aux) CPI_LAUNCH((k_elem_block4<R, PHYS, true, NT, W, STREAM, CW>), nblocks, NT, smem, s, p);
    CPI_LAUNCH((k_elem_block4<R, PHYS, false, NT, W, STREAM, CW>), nblocks, NT, smem, s, p);
  }
}

template <int NPE, int RULE, typename R, int PHYS>
static hipError_t elem_block_t(const Params &p, bool aux, int nblocks, size_t smem, hipStream_t s,
                               int part) {
  constexpr int NT = block_threads<NPE, RULE>();
  if constexpr (NPE == 4) {
    // a plan whose blocks exceed the corner words has none uploaded, and
    // its state cached
    if (!p.corner) return elem_block4_t<R, PHYS, false, false>(p, aux, nblocks, smem, s, part);
    if constexpr (PHYS != PHYS_COMPAT)
      if (p.stream_state) return elem_block4_t<R, PHYS, true, true>(p, aux, nblocks, smem, s, part);
    return elem_block4_t<R, PHYS, false, true>(p, aux, nblocks, smem, s, part);
  } else if constexpr (Quad<RULE>::n == 4) {
    constexpr int W = tet10_waves<R>();
    if constexpr (PHYS != PHYS_COMPAT) {
      if (p.stream_state) {
        if (aux)
          CPI_LAUNCH((k_elem_block10<RULE, R, PHYS, true, NT, W, true>), nblocks, NT, smem, s, p);
        CPI_LAUNCH((k_elem_block10<RULE, R, PHYS, false, NT, W, true>), nblocks, NT, smem, s, p);
      }
    }
    if (aux) CPI_LAUNCH((k_elem_block10<RULE, R, PHYS, true, NT, W, false>), nblocks, NT, smem, s, p);
    CPI_LAUNCH((k_elem_block10<RULE, R, PHYS, false, NT, W, false>), nblocks, NT, smem, s, p);
  } else {
    if (aux) CPI_LAUNCH((k_elem_block<NPE, RULE, R, PHYS, true, NT>), nblocks, NT, smem, s, p);
    CPI_LAUNCH((k_elem_block<NPE, RULE, R, PHYS, false, NT>), nblocks, NT, smem, s, p);
  }
}

template <int NPE, int RULE, typename R, int PHYS>
static hipError_t elem_direct_t(const Params &p, bool atomic, hipStream_t s) {
  constexpr int NT = NPE == 4 ? 256 : 128;
  const int grid = (p.elem_count + NT - 1) / NT;
  if (grid == 0) return hipSuccess;
  if (atomic) CPI_LAUNCH((k_elem_direct<NPE, RULE, R, PHYS, true, NT>), grid, NT, 0, s, p);
  CPI_LAUNCH((k_elem_direct<NPE, RULE, R, PHYS, false, NT>), grid, NT, 0, s, p);
}

int elem_direct_threads(int npe) { return npe == 4 ? 256 : 128; }

// (npe, rule, precision, physics) -> instantiation
template <template <int, int, typename, int> class F, typename... A>
static hipError_t dispatch(int npe, int rule, int prec, int phys, A &&...a) {
#define CPI_PHYS(NPE, RULE, R)                                         \
  switch (phys) {                                                      \
    case PHYS_COMPAT: return F<NPE, RULE, R, PHYS_COMPAT>::run(a...); \
    case PHYS_EP: return F<NPE, RULE, R, PHYS_EP>::run(a...);         \
    case PHYS_EP_MLP: return F<NPE, RULE, R, PHYS_EP_MLP>::run(a...); \
  }                                                                    \
  return hipErrorInvalidValue;
#define CPI_PREC(NPE, RULE)    \
  if (prec == 1) {             \
    CPI_PHYS(NPE, RULE, float) \
  } else {                     \
    CPI_PHYS(NPE, RULE, double) \
  }
  if (npe == 4) {
    CPI_PREC(4, RULE_CENTROID)
  }
  switch (rule) {
    case RULE_REF5: CPI_PREC(10, RULE_REF5)
    case RULE_REF4: CPI_PREC(10, RULE_REF4)
    case RULE_EXACT4: CPI_PREC(10, RULE_EXACT4)
  }
#undef CPI_PREC
#undef CPI_PHYS
  return hipErrorInvalidValue;
}

template <int NPE, typename R, int PHYS>
constexpr bool valid_combo() {
  return PHYS != PHYS_EP_MLP || (NPE == 4 && std::is_same<R, float>::value);
}

template <int NPE, int RULE, typename R, int PHYS> struct ElemBlockF {
  static hipError_t run(const Params &p, bool aux, int nb, size_t smem, hipStream_t s, int part) {
    if constexpr (!valid_combo<NPE, R, PHYS>())
      return hipErrorInvalidValue;
    else
      return elem_block_t<NPE, RULE, R, PHYS>(p, aux, nb, smem, s, part);
  }
};
template <int NPE, int RULE, typename R, int PHYS> struct ElemBlockPrepF {
  static hipError_t run(size_t smem) {
    if constexpr (!valid_combo<NPE, R, PHYS>()) {
      return hipErrorInvalidValue;
    } else {
      constexpr int NT = block_threads<NPE, RULE>();
      if constexpr (NPE == 4) {
        hipError_t e = tet4_opt_in<R, PHYS, false, true>(smem);
        if (e == hipSuccess) e = tet4_opt_in<R, PHYS, false, false>(smem);
        if constexpr (PHYS != PHYS_COMPAT)
          if (e == hipSuccess) e = tet4_opt_in<R, PHYS, true, true>(smem);
        return e;
      } else if constexpr (Quad<RULE>::n == 4) {
        constexpr int W = tet10_waves<R>();
        hipError_t e = lds_opt_in(k_elem_block10<RULE, R, PHYS, true, NT, W, false>, smem);
        if (e == hipSuccess) e = lds_opt_in(k_elem_block10<RULE, R, PHYS, false, NT, W, false>, smem);
        if constexpr (PHYS != PHYS_COMPAT) {
          if (e == hipSuccess) e = lds_opt_in(k_elem_block10<RULE, R, PHYS, true, NT, W, true>, smem);
          if (e == hipSuccess) e = lds_opt_in(k_elem_block10<RULE, R, PHYS, false, NT, W, true>, smem);
        }
        return e;
      } else {
        hipError_t e = lds_opt_in(k_elem_block<NPE, RULE, R, PHYS, true, NT>, smem);
        if (e != hipSuccess) return e;
        return lds_opt_in(k_elem_block<NPE, RULE, R, PHYS, false, NT>, smem);
      }
    }
  }
};
template <int NPE, int RULE, typename R, int PHYS> struct ElemDirectF {
  static hipError_t run(const Params &p, bool atomic, hipStream_t s) {
    if constexpr (PHYS == PHYS_EP_MLP)
      return hipErrorInvalidValue;
    else
      return elem_direct_t<NPE, RULE, R, PHYS>(p, atomic, s);
  }
};

hipError_t launch_elem_block(const Params &p, int rule, int prec, int phys, bool aux, int nblocks,
                             size_t smem, hipStream_t s, int part) {
  return dispatch<ElemBlockF>(p.npe, rule, prec, phys, p, aux, nblocks, smem, s, part);
}

hipError_t prepare_elem_block(int npe, int rule, int prec, int phys, size_t smem) {
  return dispatch<ElemBlockPrepF>(npe, rule, prec, phys, smem);
}

hipError_t launch_elem_direct(const Params &p, int rule, int prec, int phys, bool atomic,
                              hipStream_t s) {
  return dispatch<ElemDirectF>(p.npe, rule, prec, phys, p, atomic, s);
}

hipError_t launch_node_bnd(const Params &p, int prec, bool aux, hipStream_t s) {
  const int grid = std::max(1, (p.n_boundary + 255) / 256);
  if (prec == 1) {
    if (aux) CPI_LAUNCH((k_node_bnd<true, float>), grid, 256, 0, s, p);
    CPI_LAUNCH((k_node_bnd<false, float>), grid, 256, 0, s, p);
  }
  if (aux) CPI_LAUNCH((k_node_bnd<true, double>), grid, 256, 0, s, p);
  CPI_LAUNCH((k_node_bnd<false, double>), grid, 256, 0, s, p);
}

hipError_t launch_node_all(const Params &p, bool aux, hipStream_t s) {
  const int grid = (p.nn + 255) / 256;
  if (aux) CPI_LAUNCH(k_node_all<true>, grid, 256, 0, s, p);
  CPI_LAUNCH(k_node_all<false>, grid, 256, 0, s, p);
}

hipError_t launch_node_init(const Params &p, hipStream_t s) {
  CPI_LAUNCH(k_node_init, (p.nn + 255) / 256, 256, 0, s, p);
}

hipError_t launch_energy_reduce(const Params &p, hipStream_t s) {
  hipLaunchKernelGGL(k_energy_reduce1, dim3(kEredBlocks), dim3(256), 0, s, p);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  CPI_LAUNCH(k_energy_reduce2, 1, 64, 0, s, p);
}

int energy_reduce_blocks() { return kEredBlocks; }

template <int NPE, int RULE>
static hipError_t mass_rule_t(const Params &p, int scatter, int kind, double rho, int nblocks,
                              hipStream_t s) {
  if (scatter == 0) {
    constexpr int NT = NPE == 4 ? 512 : 128;
    const size_t smem = 8 * (3 * (size_t)p.ml + (size_t)p.max_el * NPE) + lds_off_bytes(p.kmax);
    if (smem > 65536) return hipErrorInvalidValue;  // sized by the plan to stay below
    hipLaunchKernelGGL((k_mass_block<NPE, RULE, NT>), dim3(nblocks), dim3(NT), smem, s, p, kind,
                       rho);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || p.n_boundary == 0) return e;
    CPI_LAUNCH(k_mass_bnd, (p.n_boundary + 255) / 256, 256, 0, s, p);
  }
  const int grid = (p.elem_count + 255) / 256;
  if (grid == 0) return hipSuccess;
  if (scatter == 2) CPI_LAUNCH((k_mass_direct<NPE, RULE, true>), grid, 256, 0, s, p, kind, rho);
  CPI_LAUNCH((k_mass_direct<NPE, RULE, false>), grid, 256, 0, s, p, kind, rho);
}

hipError_t launch_mass(const Params &p, int rule, int scatter, int kind, double rho, int nblocks,
                       hipStream_t s) {
  if (p.npe == 4) {
    switch (rule) {
      case RULE_REF4: return mass_rule_t<4, RULE_REF4>(p, scatter, kind, rho, nblocks, s);
      case RULE_EXACT4: return mass_rule_t<4, RULE_EXACT4>(p, scatter, kind, rho, nblocks, s);
      case RULE_CENTROID: return mass_rule_t<4, RULE_CENTROID>(p, scatter, kind, rho, nblocks, s);
      default: return mass_rule_t<4, RULE_REF5>(p, scatter, kind, rho, nblocks, s);
    }
  }
  switch (rule) {
    case RULE_REF4: return mass_rule_t<10, RULE_REF4>(p, scatter, kind, rho, nblocks, s);
    case RULE_EXACT4: return mass_rule_t<10, RULE_EXACT4>(p, scatter, kind, rho, nblocks, s);
    default: return mass_rule_t<10, RULE_REF5>(p, scatter, kind, rho, nblocks, s);
  }
}

hipError_t launch_elem_dt(const Params &p, double *out, hipStream_t s) {
  const int grid = (p.ne + 255) / 256;
  if (p.npe == 4) CPI_LAUNCH(k_elem_dt<4>, grid, 256, 0, s, p, out);
  CPI_LAUNCH(k_elem_dt<10>, grid, 256, 0, s, p, out);
}

hipError_t launch_stress_elem(const Params &p, double *out, hipStream_t s) {
  const int grid = (p.ne + 255) / 256;
  if (p.npe == 4) CPI_LAUNCH(k_stress_elem<4>, grid, 256, 0, s, p, out);
  CPI_LAUNCH(k_stress_elem<10>, grid, 256, 0, s, p, out);
}

hipError_t launch_debug_strain(const Params &p, int rule, double alpha, int dcase, double nu,
                               double *out, hipStream_t s) {
  const int grid = (p.ne + 255) / 256;
#define CPI_DS(NPE, RULE) \
  CPI_LAUNCH((k_debug_strain<NPE, RULE>), grid, 256, 0, s, p, alpha, dcase, nu, out)
  if (p.npe == 4) {
    switch (rule) {
      case RULE_REF4: CPI_DS(4, RULE_REF4);
      case RULE_EXACT4: CPI_DS(4, RULE_EXACT4);
      case RULE_CENTROID: CPI_DS(4, RULE_CENTROID);
      default: CPI_DS(4, RULE_REF5);
    }
  }
  switch (rule) {
    case RULE_REF4: CPI_DS(10, RULE_REF4);
    case RULE_EXACT4: CPI_DS(10, RULE_EXACT4);
    default: CPI_DS(10, RULE_REF5);
  }
#undef CPI_DS
}

hipError_t launch_mlp_rows(const MlpParams *P, uint16_t *rows, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int tiles = (n + 15) / 16;
  const int grid = std::min(kMlpQueueGrid, (tiles + 3) / 4);
  CPI_LAUNCH((k_mlp_queue<kMlpQueueThreads>), grid, kMlpQueueThreads, 0, s, P, rows,
             (const unsigned long long *)nullptr, n);
}

hipError_t launch_stream_copy(const void *src, void *dst, long n4, hipStream_t s) {
  const int grid = (int)std::min<long>((n4 + 255) / 256, 8L * 1024);
  CPI_LAUNCH(k_stream_copy, grid, 256, 0, s, (const Node4 *)src, (Node4 *)dst, n4);
}

}  // namespace cpi
