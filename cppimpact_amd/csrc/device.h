// Device-side data structures and element/node physics shared by the
// kernels of libcppimpact_hip.so (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <limits>
#include <type_traits>

namespace cpi {

constexpr int kMaxQ = 5;

// energy partial row: one per workgroup of the element and node kernels
enum : int {
  EP_WINT = 0,  // cumulative internal work
  EP_WPL,       // cumulative plastic dissipation
  EP_MINDT,     // stable dt of this step's configuration (min)
  EP_BAD,       // cumulative count of elements with detJ <= 0
  EP_KE,        // kinetic energy of the nodes updated by the workgroup
  EP_PEN,       // penalty energy
  EP_FC,        // contact force
  EP_WGRAV,     // cumulative gravity work
  EP_N
};

struct Ctl {
  double dt;
  int64_t step;    // completed loop iterations
  int64_t nrows;   // energy rows written to the history ring
  uint32_t flags;  // bit 0: non-finite energy, bit 1: inverted element,
                   // bit 2: dt above the stable limit of the configuration
  uint32_t pad;
  double e0;       // KE + Epen of the initial state
  double row[8];   // latest energy row (CPI_E_* order)
  double dt_contact;  // wall-spring limit 2 sqrt(m_min / k) (fixed: mass is)
};

struct MlpParams;  // surrogate, see mlp.h

struct alignas(32) Node4 {
  double x, y, z, w;
};

// Packed (x, y, z) records, 24 B in FP64 and 12 B in FP32: velocities and
// the blocks' partial forces on shared nodes (a padded fourth word would add
// a third to the bytes each step reads and writes of them)
struct alignas(8) Vec3 {
  double x, y, z;
};
struct alignas(4) Vec3f {
  float x, y, z;
};

struct Params {
  int nn, ne, npe, nq;
  // material (D per tetrahedral.h:486-512: c11 = (1-nu) k, c12 = nu k,
  // c44 = (1-2nu)/2 k with k = E/((1+nu)(1-2nu)))
  double c11, c12, c44, mu, sy0, H, wave_speed;
  // wall (wall.h:37-44)
  int wall_on, wall_dim, wall_normal, pad0;
  double wall_loc, wall_k, wall_nk;
  const uint8_t *contact_mask;  // slave-only contact, or null for all nodes
  const uint8_t *fixed_mask;    // held nodes (fixed set), or null for none
  double grav[3];
  // quadrature points (tet10) and the weight sum (tet4)
  double qw[kMaxQ];
  double qpt[kMaxQ][3];
  double wsum;
  // nodal state, device node order: one record per node per field, so a
  // gathered node's x (or v) costs one or two 32-byte sectors
  Vec3 *pos;             // x
  Vec3 *vv;              // v_{n+1/2}
  double *m;             // lumped mass (read by the node's owner only)
  double *f[3];          // assembled force (colour / atomic scatter)
  Node4 *acc;            // a_n                   written when store_aux
  Node4 *du;             // du of the step        written when store_aux
  Node4 *xprev;          // x_n before the step   written when store_aux
  Node4 *X0;             // reference positions
  const int32_t *conn;   // [npe][ne] device node ids (colour / atomic)
  void *sig;             // [nq][6][ne] stress state (ep)
  void *epbar;           // [nq][ne]
  // tet4 and per-point tet10 block kernels, exact return: one bit per
  // stress point, set once it has yielded (ep > 0), ywords 64-bit words per
  // block (the point of thread t of block b at word b*ywords + t/64, bit
  // t%64); ep of the others is 0 and is not read.  Null: ep is always read.
  unsigned long long *yielded;
  int ywords;
  // 1: the tet4 and per-point tet10 block kernels read and write the stress
  // state with non-temporal accesses (a working set larger than the cache)
  int stream_state;
  // workgroup plan (LDS-assembled scatter)
  const int32_t *meta;
  const uint16_t *conn_local;
  const uint16_t *csr_pos;   // [E][npe]: corner -> force position in its block
  // [E][npe]: corner as l | k << 9, its force at off[k] + l (the per-point
  // tet10 block kernel; the tet4 ones when the plan fits the words, else
  // null and they read conn_local and csr_pos)
  const uint16_t *corner;
  const int32_t *bnd_nodes;
  const int32_t *bnd_slot;
  const int32_t *slot_ptr;
  Vec3 *partial;        // one force partial per slot (Vec3f in FP32)
  // per-block node tables at a fixed stride ml (local node j of block b at
  // b*ml + j), read without the block's meta: device node id (-1 past the
  // block's nodes), partial slot (-1 for interior nodes), contribution
  // count; and kmax+1 slot offsets off[k] per block (jagged-diagonal corner
  // force slots: local node j's k-th contribution at off[k] + j)
  const int32_t *nid_pad;
  const int32_t *slot_pad;
  const uint16_t *deg_pad;
  const uint16_t *off_pad;
  int ml, kmax;
  int n_interior, n_boundary, max_loc, max_el;
  // energy partial rows and control block
  double *epart;         // one row of EP_N doubles per wave of the step kernels
  double *ered;          // first-stage sums of the rows
  int epart_base;        // first row of this launch
  int n_epart;
  Ctl *ctl;
  double *hist;          // [hist_cap][8]
  int hist_cap;
  int store_aux;
  int elem_base;         // colour / atomic: first device element of launch
  int elem_count;
  const MlpParams *mlp;  // null unless the surrogate is enabled
  double mlp_tol;        // fallback threshold on |f_yield| / sy0
  unsigned long long *mlp_stats;  // [nblocks][2] surrogate accepted / fell back
  // surrogate mode (block scatter): the elastic pass defers a block with a
  // yielding element, lists it and queues the network inputs of its
  // yielding points; mlp_ctr = (deferred blocks << 32) | queued points of
  // the step, zeroed by the boundary-node kernel at the end of the step
  unsigned long long *mlp_ctr;
  int2 *mlp_list;        // [nblocks] (block, its first queue row)
  uint16_t *mlp_rows;    // [ne + 16] rows of 32 bytes (see mlp_infer.h)
  int nblocks;           // element blocks of the block scatter
  float *samples;        // return-mapping sample buffer (optional)
  int *sample_count;
  int sample_cap;
  // a yielding point is recorded when the 32-bit hash of (seed, step,
  // element, point) is below sample_keep (2^32: every point)
  unsigned long long sample_keep;
  unsigned long long sample_seed;
};

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

// Element blocks in chunks of kXcdChunk consecutive (Hilbert-ordered, so
// neighbouring) blocks, the chunks dealt round-robin to the 8 XCDs: the
// blocks of a chunk share nodes in one L2, and a hot region spreads over
// every XCD instead of landing on the few whose contiguous range covers it.
// The surrogate kernel uses it: its plastic zone at the wall costs several
// times an elastic block.  Hardware workgroup hw runs on XCD hw & 7 as its
// (hw >> 3)-th; blocks past the last whole round of chunks keep their
// index.
constexpr int kXcdChunk = 16;
__device__ __forceinline__ int xcd_chunked(int hw, int n) {
  const int m = n / (8 * kXcdChunk) * (8 * kXcdChunk);
  if (hw >= m) return hw;
  const int x = hw & 7, k = hw >> 3;
  return ((k / kXcdChunk) * 8 + x) * kXcdChunk + k % kXcdChunk;
}

// hardware dispatches workgroups round-robin over the 8 XCDs; give each XCD
// a contiguous range of logical blocks so neighbouring blocks, which share
// nodes, hit the same L2
__device__ __forceinline__ int xcd_remap(int hw, int n) {
  const int q = n >> 3, r = n & 7;
  const int xcd = hw & 7, k = hw >> 3;
  return xcd * q + min(xcd, r) + k;
}

// DPP lane moves (the source lane must be active).  CTRL: 0xB1 lane ^ 1 and
// 0x4E lane ^ 2 (quad_perm), 0x141 the mirror lane within each 8 lanes,
// 0x140 the mirror lane within each row of 16.
enum : int { kDppXor1 = 0xB1, kDppXor2 = 0x4E, kDppHalfMirror = 0x141, kDppMirror = 0x140 };
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// lane L's value, the same on every lane
__device__ __forceinline__ double lane_value(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// sum over each group of four lanes, (v0 + v1) + (v2 + v3) on every lane of
// the group (addition commutes, so all four get the same bits)
template <typename R>
__device__ __forceinline__ R quad_sum(R v) {
  v = v + dpp_mov<kDppXor1>(v);
  return v + dpp_mov<kDppXor2>(v);
}

// op over each row of 16 lanes with DPP moves (xor 1, xor 2, half-row
// mirror, row mirror); op commutes, so every lane of a row ends with the
// same bits
template <class Op>
__device__ __forceinline__ double row_reduce(double v, const Op &op) {
  v = op(v, dpp_mov<kDppXor1>(v));
  v = op(v, dpp_mov<kDppXor2>(v));
  v = op(v, dpp_mov<kDppHalfMirror>(v));
  return op(v, dpp_mov<kDppMirror>(v));
}

template <typename R>
__device__ __forceinline__ R inv3(const R J[9], R Ji[9]) {
  R c0 = J[4] * J[8] - J[5] * J[7];
  R c1 = J[5] * J[6] - J[3] * J[8];
  R c2 = J[3] * J[7] - J[4] * J[6];
  R det = J[0] * c0 + J[1] * c1 + J[2] * c2;
  R r = R(1) / det;
  Ji[0] = c0 * r;
  Ji[1] = (J[2] * J[7] - J[1] * J[8]) * r;
  Ji[2] = (J[1] * J[5] - J[2] * J[4]) * r;
  Ji[3] = c1 * r;
  Ji[4] = (J[0] * J[8] - J[2] * J[6]) * r;
  Ji[5] = (J[2] * J[3] - J[0] * J[5]) * r;
  Ji[6] = c2 * r;
  Ji[7] = (J[1] * J[6] - J[0] * J[7]) * r;
  Ji[8] = (J[0] * J[4] - J[1] * J[3]) * r;
  return det;
}

// quadratic basis gradients at (x,y,z), tetrahedral.h:199-250 layout
template <typename T>
__device__ __forceinline__ void tet10_grad(T x, T y, T z, T G[30]) {
  G[0] = 4 * x - 1; G[1] = 0.0; G[2] = 0.0;
  G[3] = 0.0; G[4] = 4 * y - 1; G[5] = 0.0;
  G[6] = 4 * x + 4 * y + 4 * z - 3;
  G[7] = 4 * y + 4 * x + 4 * z - 3;
  G[8] = 4 * z + 4 * x + 4 * y - 3;
  G[9] = 0.0; G[10] = 0.0; G[11] = 4 * z - 1;
  G[12] = 4 * y; G[13] = 4 * x; G[14] = 0.0;
  G[15] = -4 * y; G[16] = -8 * y - 4 * x - 4 * z + 4; G[17] = -4 * y;
  G[18] = -8 * x - 4 * y - 4 * z + 4; G[19] = -4 * x; G[20] = -4 * x;
  G[21] = 4 * z; G[22] = 0.0; G[23] = 4 * x;
  G[24] = 0.0; G[25] = 4 * z; G[26] = 4 * y;
  G[27] = -4 * z; G[28] = -4 * z; G[29] = -8 * z - 4 * x - 4 * y + 4;
}

// energies accumulated by one thread
struct EAcc {
  double wint = 0.0, wpl = 0.0, mindt = 1e300, bad = 0.0;
  double ke = 0.0, pen = 0.0, fc = 0.0, wgrav = 0.0;
};

// yield function of a trial stress: |dev s| - sqrt(2/3) (sy0 + H ep), the
// same arithmetic as radial_return's test
template <typename R>
__device__ __forceinline__ R yield_overshoot(const Params &p, const R s[6], R ep) {
  const R third = R(1) / R(3);
  R pm = (s[0] + s[1] + s[2]) * third;
  R d0 = s[0] - pm, d1 = s[1] - pm, d2 = s[2] - pm;
  R nrm = sqrt(d0 * d0 + d1 * d1 + d2 * d2 + R(2) * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
  const R k23 = R(0.81649658092772603273);  // sqrt(2/3)
  return nrm - k23 * (R(p.sy0) + R(p.H) * ep);
}

// A trial stress whose squared deviator norm q2 lies below the squared
// yield radius by 64 epsilon: far more than the rounding of the square
// root, the squares and the radius (a few epsilon), so the test
// sqrt(q2) - radius > 0 leaves every such stress elastic as well, and
// skipping the square root for it changes no decision and no bit.  FP64
// only: an FP64 square root is a dozen instructions at the FP64 rate, an
// FP32 one about one, which the test would cost as much as it saves.
template <typename R>
__device__ __forceinline__ bool surely_elastic(const Params &p, R q2, R ep) {
  if constexpr (!std::is_same<R, double>::value) {
    return false;
  } else {
    const R k23 = R(0.81649658092772603273);  // sqrt(2/3)
    const R rad = k23 * (R(p.sy0) + R(p.H) * ep);
    return q2 <= rad * rad * (R(1) - R(64) * std::numeric_limits<R>::epsilon());
  }
}

// yield_overshoot(p, s, ep) > 0, the square root taken only near or
// beyond the yield surface
template <typename R>
__device__ __forceinline__ bool trial_yields(const Params &p, const R s[6], R ep) {
  const R third = R(1) / R(3);
  R pm = (s[0] + s[1] + s[2]) * third;
  R d0 = s[0] - pm, d1 = s[1] - pm, d2 = s[2] - pm;
  const R q2 = d0 * d0 + d1 * d1 + d2 * d2 + R(2) * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
  if (surely_elastic<R>(p, q2, ep)) return false;
  const R k23 = R(0.81649658092772603273);  // sqrt(2/3)
  return sqrt(q2) - k23 * (R(p.sy0) + R(p.H) * ep) > R(0);
}

// J2 return map, linear isotropic hardening (Simo & Hughes 1998 Box 3.2
// restricted to isotropic hardening); s holds the trial stress on entry.
template <typename R>
__device__ __forceinline__ R radial_return(const Params &p, R s[6], R &ep) {
  const R third = R(1) / R(3);
  R pm = (s[0] + s[1] + s[2]) * third;
  R d0 = s[0] - pm, d1 = s[1] - pm, d2 = s[2] - pm;
  const R q2 = d0 * d0 + d1 * d1 + d2 * d2 + R(2) * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]);
  if (surely_elastic<R>(p, q2, ep)) return R(0);
  R nrm = sqrt(q2);
  const R k23 = R(0.81649658092772603273);  // sqrt(2/3)
  R fy = nrm - k23 * (R(p.sy0) + R(p.H) * ep);
  if (!(fy > R(0))) return R(0);
  R dg = fy / (R(2) * R(p.mu) + R(2.0 / 3.0) * R(p.H));
  R sc = R(2) * R(p.mu) * dg / nrm;
  s[0] -= sc * d0;
  s[1] -= sc * d1;
  s[2] -= sc * d2;
  s[3] -= sc * s[3];
  s[4] -= sc * s[4];
  s[5] -= sc * s[5];
  ep += k23 * dg;
  return dg;
}

}  // namespace cpi
