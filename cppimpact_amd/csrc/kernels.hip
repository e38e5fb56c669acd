// HIP kernels of the explicit-dynamics step for gfx950 and their launchers.
//
// One loop iteration of Dynamics::solve (dynamics.h:471-511) is, with the
// LDS-assembled scatter (the default):
//   k_elem_block  one workgroup per element block: stage the block's nodes
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
#include <hip/hip_runtime.h>

#include "device.h"
#include "elem.h"
#include "kernels.h"
#include "mlp_infer.h"

namespace cpi {

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

// Deterministic block reduction of the energy accumulators; the result is
// valid in thread 0.  red must hold 8 * (blockDim/64) doubles.
__device__ __forceinline__ void block_reduce(EAcc &ea, double *red, double out[EP_N]) {
  double v[EP_N] = {ea.wint, ea.wpl, ea.mindt, ea.bad, ea.ke, ea.pen, ea.fc, ea.wgrav};
#pragma unroll
  for (int k = 0; k < EP_N; k++) v[k] = (k == EP_MINDT) ? wave_min(v[k]) : wave_sum(v[k]);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < EP_N; k++) red[w * EP_N + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < EP_N; k++) out[k] = red[k];
    for (int i = 1; i < nw; i++)
      for (int k = 0; k < EP_N; k++)
        out[k] = (k == EP_MINDT) ? fmin(out[k], red[i * EP_N + k]) : out[k] + red[i * EP_N + k];
  }
}

// cumulative terms accumulate in the row, instantaneous ones overwrite it
__device__ __forceinline__ void write_row(const Params &p, int row, const double v[EP_N]) {
  double *r = p.epart + (long)row * EP_N;
  r[EP_WINT] += v[EP_WINT];
  r[EP_WPL] += v[EP_WPL];
  r[EP_BAD] += v[EP_BAD];
  r[EP_WGRAV] += v[EP_WGRAV];
  r[EP_MINDT] = v[EP_MINDT];
  r[EP_KE] = v[EP_KE];
  r[EP_PEN] = v[EP_PEN];
  r[EP_FC] = v[EP_FC];
}

// ---------------------------------------------------------------------------
// node update: a = -f/m + wall + g; x += du; v += dt a  (dynamics.h:497-503,
// wall.h:37-44, dynamics_kernels.h:150-168).  The wall sees x + du = x_{n+1}.
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
  if (p.wall_on && (!p.contact_mask || p.contact_mask[g])) {
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
  double ke = 0.0, wg = 0.0;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    a[c] += p.grav[c];
    double vn = v[c] + dt * a[c];
    double vi = vn - 0.5 * dt * a[c];
    ke += vi * vi;
    wg += p.grav[c] * du[c];
    p.x[c][g] = x[c] + du[c];
    p.v[c][g] = vn;
    if (AUX) {
      p.acc[c][g] = a[c];
      p.du[c][g] = du[c];
      p.xprev[c][g] = x[c];
    }
  }
  ea.ke += 0.5 * m * ke;
  ea.wgrav += m * wg;
}

// ---------------------------------------------------------------------------
// LDS-assembled element pass
// ---------------------------------------------------------------------------

// LDS: node records [max_loc][8] doubles (x, v, -, -), then the element
// corner forces [max_el*NPE][3] (R), then the reduction scratch.
template <int NPE, int RULE, typename R, int PHYS, bool AUX, int NT>
__global__ void __launch_bounds__(NT) k_elem_block(Params p) {
  extern __shared__ double smem[];
  const int b = xcd_remap(blockIdx.x, gridDim.x);
  const int32_t *mt = p.meta + 8 * b;
  const int es = mt[0], nel = mt[1], is = mt[2], nint = mt[3], boff = mt[4], nbnd = mt[5],
            coff = mt[6], loff = mt[7];
  const int nloc = nint + nbnd;
  double *lnode = smem;
  R *lf = (R *)(smem + 8 * p.max_loc);
  double *red = smem + 8 * p.max_loc + (sizeof(R) * 3 * NPE * p.max_el + 7) / 8;
  const double dt = p.ctl->dt;

  for (int j = threadIdx.x; j < nloc; j += NT) {
    int g = j < nint ? is + j : p.bnd_nodes[boff + j - nint];
    double *d = lnode + 8 * j;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      d[c] = p.x[c][g];
      d[3 + c] = p.v[c][g];
    }
  }
  __syncthreads();

  EAcc ea;
  for (int t = threadIdx.x; t < nel; t += NT) {
    const int e = es + t;
    double xs[NPE][3], vs[NPE][3];
#pragma unroll
    for (int k = 0; k < NPE; k++) {
      const int l = p.conn_local[(long)e * NPE + k];
      const double *d = lnode + 8 * l;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        xs[k][c] = d[c];
        vs[k][c] = d[3 + c];
      }
    }
    R f[NPE][3];
    elem_force<NPE, RULE, R, PHYS>(p, xs, vs, dt, e, f, ea);
#pragma unroll
    for (int k = 0; k < NPE; k++)
#pragma unroll
      for (int c = 0; c < 3; c++) lf[(t * NPE + k) * 3 + c] = f[k][c];
  }
  __syncthreads();

  const uint16_t *cp = p.csr_ptr + loff;
  const uint16_t *ci = p.csr_idx + coff;
  for (int j = threadIdx.x; j < nloc; j += NT) {
    double f[3] = {0.0, 0.0, 0.0};
    const int s0 = cp[j], s1 = cp[j + 1];
    for (int s = s0; s < s1; s++) {
      const R *src = lf + 3 * ci[s];
      f[0] += double(src[0]);
      f[1] += double(src[1]);
      f[2] += double(src[2]);
    }
    if (j < nint) {
      const double *d = lnode + 8 * j;
      const int g = is + j;
      double x[3] = {d[0], d[1], d[2]}, v[3] = {d[3], d[4], d[5]};
      node_update<AUX>(p, g, f, x, v, p.m[g], dt, ea);
    } else {
      double *dst = p.partial + 3L * p.bnd_slot[boff + j - nint];
      dst[0] = f[0];
      dst[1] = f[1];
      dst[2] = f[2];
    }
  }
  double row[EP_N];
  block_reduce(ea, red, row);
  if (threadIdx.x == 0) write_row(p, p.epart_base + b, row);
}

template <bool AUX>
__global__ void __launch_bounds__(256) k_node_bnd(Params p) {
  __shared__ double red[4 * EP_N];
  const int i = blockIdx.x * 256 + threadIdx.x;
  const double dt = p.ctl->dt;
  EAcc ea;
  if (i < p.n_boundary) {
    const int g = p.n_interior + i;
    double f[3] = {0.0, 0.0, 0.0};
    for (int s = p.slot_ptr[i]; s < p.slot_ptr[i + 1]; s++) {
      const double *src = p.partial + 3L * s;
      f[0] += src[0];
      f[1] += src[1];
      f[2] += src[2];
    }
    double x[3], v[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      x[c] = p.x[c][g];
      v[c] = p.v[c][g];
    }
    node_update<AUX>(p, g, f, x, v, p.m[g], dt, ea);
  }
  double row[EP_N];
  block_reduce(ea, red, row);
  if (threadIdx.x == 0) {
    write_row(p, p.epart_base + blockIdx.x, row);
    if (blockIdx.x == 0) p.ctl->step += 1;
  }
}

// ---------------------------------------------------------------------------
// colour / atomic element pass: one thread per element, nodes from HBM
// ---------------------------------------------------------------------------
template <int NPE, int RULE, typename R, int PHYS, bool ATOMIC, int NT>
__global__ void __launch_bounds__(NT) k_elem_direct(Params p) {
  __shared__ double red[(NT / 64) * EP_N];
  const int t = blockIdx.x * NT + threadIdx.x;
  const double dt = p.ctl->dt;
  EAcc ea;
  if (t < p.elem_count) {
    const int e = p.elem_base + t;
    int nd[NPE];
    double xs[NPE][3], vs[NPE][3];
#pragma unroll
    for (int k = 0; k < NPE; k++) {
      nd[k] = p.conn[(long)k * p.ne + e];
#pragma unroll
      for (int c = 0; c < 3; c++) {
        xs[k][c] = p.x[c][nd[k]];
        vs[k][c] = p.v[c][nd[k]];
      }
    }
    R f[NPE][3];
    elem_force<NPE, RULE, R, PHYS>(p, xs, vs, dt, e, f, ea);
#pragma unroll
    for (int k = 0; k < NPE; k++)
#pragma unroll
      for (int c = 0; c < 3; c++) {
        if (ATOMIC)
          unsafeAtomicAdd(&p.f[c][nd[k]], double(f[k][c]));
        else
          p.f[c][nd[k]] += double(f[k][c]);
      }
  }
  double row[EP_N];
  block_reduce(ea, red, row);
  if (threadIdx.x == 0) write_row(p, p.epart_base + blockIdx.x, row);
}

template <bool AUX>
__global__ void __launch_bounds__(256) k_node_all(Params p) {
  __shared__ double red[4 * EP_N];
  const int g = blockIdx.x * 256 + threadIdx.x;
  const double dt = p.ctl->dt;
  EAcc ea;
  if (g < p.nn) {
    double f[3], x[3], v[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      f[c] = p.f[c][g];
      p.f[c][g] = 0.0;
      x[c] = p.x[c][g];
      v[c] = p.v[c][g];
    }
    node_update<AUX>(p, g, f, x, v, p.m[g], dt, ea);
  }
  double row[EP_N];
  block_reduce(ea, red, row);
  if (threadIdx.x == 0) {
    write_row(p, p.epart_base + blockIdx.x, row);
    if (blockIdx.x == 0) p.ctl->step += 1;
  }
}

// initial update with du = 0 (f_int = 0) and the half-kick
// (dynamics.h:454-463): a0 = wall(X0) + g; v += dt/2 a0
__global__ void __launch_bounds__(256) k_node_init(Params p) {
  __shared__ double red[4 * EP_N];
  const int g = blockIdx.x * 256 + threadIdx.x;
  const double dt = p.ctl->dt;
  EAcc ea;
  if (g < p.nn) {
    const double m = p.m[g];
    double a[3] = {0.0, 0.0, 0.0}, ke = 0.0;
    if (p.wall_on && (!p.contact_mask || p.contact_mask[g])) {
#pragma unroll
      for (int c = 0; c < 3; c++)
        if (c == p.wall_dim) {
          double d = (p.x[c][g] - p.wall_loc) * p.wall_normal;
          if (d < 0.0) {
            a[c] += -1 * (1 / m) * d * p.wall_nk;
            ea.pen += 0.5 * p.wall_k * d * d;
            ea.fc += -p.wall_k * d;
          }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
      a[c] += p.grav[c];
      double v = p.v[c][g];
      ke += v * v;
      p.v[c][g] = v + 0.5 * dt * a[c];
      p.acc[c][g] = a[c];
      p.du[c][g] = 0.0;
      p.xprev[c][g] = p.x[c][g];
    }
    ea.ke = 0.5 * m * ke;
  }
  double row[EP_N];
  block_reduce(ea, red, row);
  if (threadIdx.x == 0) write_row(p, blockIdx.x, row);
}

// sums the partial rows in a fixed order into the energy row of this step
__global__ void __launch_bounds__(256) k_energy_reduce(Params p) {
  __shared__ double red[256 * EP_N];
  double v[EP_N] = {0, 0, 1e300, 0, 0, 0, 0, 0};
  for (int r = threadIdx.x; r < p.n_epart; r += 256) {
    const double *src = p.epart + (long)r * EP_N;
#pragma unroll
    for (int k = 0; k < EP_N; k++) v[k] = (k == EP_MINDT) ? fmin(v[k], src[k]) : v[k] + src[k];
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
  if (threadIdx.x == 0) {
    Ctl *c = p.ctl;
    const double *s = red;
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
    if (p.hist && p.hist_cap > 0) {
      double *h = p.hist + (c->nrows % p.hist_cap) * 8;
      for (int k = 0; k < 8; k++) h[k] = row[k];
      c->nrows += 1;
    }
  }
}

// ---------------------------------------------------------------------------
// lumped mass
// ---------------------------------------------------------------------------
template <int NPE, int RULE, int NT>
__global__ void __launch_bounds__(NT) k_mass_block(Params p, int kind, double rho, double *m) {
  extern __shared__ double smem[];
  const int b = blockIdx.x;
  const int32_t *mt = p.meta + 8 * b;
  const int es = mt[0], nel = mt[1], is = mt[2], nint = mt[3], boff = mt[4], nbnd = mt[5],
            coff = mt[6], loff = mt[7];
  const int nloc = nint + nbnd;
  double *lnode = smem;
  double *lm = smem + 4 * p.max_loc;
  for (int j = threadIdx.x; j < nloc; j += NT) {
    int g = j < nint ? is + j : p.bnd_nodes[boff + j - nint];
#pragma unroll
    for (int c = 0; c < 3; c++) lnode[4 * j + c] = p.x[c][g];
  }
  __syncthreads();
  for (int t = threadIdx.x; t < nel; t += NT) {
    const int e = es + t;
    double xs[NPE][3], me[NPE];
#pragma unroll
    for (int k = 0; k < NPE; k++) {
      const int l = p.conn_local[(long)e * NPE + k];
#pragma unroll
      for (int c = 0; c < 3; c++) xs[k][c] = lnode[4 * l + c];
    }
    elem_mass<NPE, RULE>(p, xs, kind, rho, me);
#pragma unroll
    for (int k = 0; k < NPE; k++) lm[t * NPE + k] = me[k];
  }
  __syncthreads();
  const uint16_t *cp = p.csr_ptr + loff;
  const uint16_t *ci = p.csr_idx + coff;
  for (int j = threadIdx.x; j < nloc; j += NT) {
    double s = 0.0;
    for (int q = cp[j]; q < cp[j + 1]; q++) s += lm[ci[q]];
    if (j < nint)
      m[is + j] = s;
    else
      p.partial[3L * p.bnd_slot[boff + j - nint]] = s;
  }
}

__global__ void __launch_bounds__(256) k_mass_bnd(Params p, double *m) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= p.n_boundary) return;
  double s = 0.0;
  for (int q = p.slot_ptr[i]; q < p.slot_ptr[i + 1]; q++) s += p.partial[3L * q];
  m[p.n_interior + i] = s;
}

template <int NPE, int RULE, bool ATOMIC>
__global__ void __launch_bounds__(256) k_mass_direct(Params p, int kind, double rho, double *m) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  if (t >= p.elem_count) return;
  const int e = p.elem_base + t;
  int nd[NPE];
  double xs[NPE][3], me[NPE];
#pragma unroll
  for (int k = 0; k < NPE; k++) {
    nd[k] = p.conn[(long)k * p.ne + e];
#pragma unroll
    for (int c = 0; c < 3; c++) xs[k][c] = p.x[c][nd[k]];
  }
  elem_mass<NPE, RULE>(p, xs, kind, rho, me);
#pragma unroll
  for (int k = 0; k < NPE; k++) {
    if (ATOMIC)
      unsafeAtomicAdd(&m[nd[k]], me[k]);
    else
      m[nd[k]] += me[k];
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
      const int n = p.conn[(long)k * p.ne + e];
#pragma unroll
      for (int c = 0; c < 3; c++) xs[k][c] = p.x[c][n];
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
    d = corner_dt<double>(e1, e2, e3, v6, p.wave_speed, NPE == 4 ? 1.0 : kTet10LengthFactor);
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
__global__ void __launch_bounds__(256) k_stress_elem(Params p, const double *X0x,
                                                      const double *X0y, const double *X0z,
                                                      double *out) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= p.ne) return;
  const double *X0[3] = {X0x, X0y, X0z};
  double xs[NPE][3], u[NPE][3];
#pragma unroll
  for (int k = 0; k < NPE; k++) {
    const int n = p.conn[(long)k * p.ne + e];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      xs[k][c] = p.xprev[c][n];
      u[k][c] = p.du[c][n] + p.xprev[c][n] - X0[c][n];
    }
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
// known-answer strain energy and volume (dynamics.h:303-412,
// analysis.h:384-477): W_e = 1/2 sum_q w detJ de.D.de, V_e = sum_q w detJ
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
      const int n = p.conn[(long)k * p.ne + e];
#pragma unroll
      for (int c = 0; c < 3; c++) xs[k][c] = p.x[c][n];
      const double x = xs[k][0];
      const double s = dcase == 0 ? alpha * x : alpha * 0.5 * x * x;
      u[k][0] = s;
      u[k][1] = -s * nu;
      u[k][2] = -s * nu;
    }
    constexpr int NQ = NPE == 4 ? Quad<RULE>::n : Quad<RULE>::n;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
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
// launchers
// ---------------------------------------------------------------------------

#define CPI_LAUNCH(kern, grid, block, smem, stream, ...)                       \
  do {                                                                          \
    hipLaunchKernelGGL((kern), dim3(grid), dim3(block), (smem), (stream), __VA_ARGS__); \
    return hipGetLastError();                                                   \
  } while (0)

template <int NPE> constexpr int block_threads() { return NPE == 4 ? 512 : 128; }

size_t elem_block_smem(int npe, int prec, int max_loc, int max_el) {
  const size_t rsz = prec == 1 ? 4 : 8;
  const int nt = npe == 4 ? block_threads<4>() : block_threads<10>();
  return 8 * 8 * (size_t)max_loc + ((rsz * 3 * npe * max_el + 7) / 8) * 8 + 8 * EP_N * (nt / 64);
}

int elem_block_threads(int npe) { return npe == 4 ? block_threads<4>() : block_threads<10>(); }

// dynamic LDS above 64 KiB needs an explicit opt-in per kernel
template <typename K>
static hipError_t lds_opt_in(K kernel, size_t smem) {
  if (smem <= 65536) return hipSuccess;
  return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                             hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
}

// the LDS opt-in (prepare_elem_block) is done once at set-up, outside any
// stream capture
template <int NPE, int RULE, typename R, int PHYS>
static hipError_t elem_block_t(const Params &p, bool aux, int nblocks, size_t smem, hipStream_t s) {
  constexpr int NT = block_threads<NPE>();
  if (aux) CPI_LAUNCH((k_elem_block<NPE, RULE, R, PHYS, true, NT>), nblocks, NT, smem, s, p);
  CPI_LAUNCH((k_elem_block<NPE, RULE, R, PHYS, false, NT>), nblocks, NT, smem, s, p);
}

template <int NPE, int RULE, typename R, int PHYS>
static hipError_t elem_direct_t(const Params &p, bool atomic, hipStream_t s) {
  constexpr int NT = NPE == 4 ? 256 : 64;
  const int grid = (p.elem_count + NT - 1) / NT;
  if (grid == 0) return hipSuccess;
  if (atomic) CPI_LAUNCH((k_elem_direct<NPE, RULE, R, PHYS, true, NT>), grid, NT, 0, s, p);
  CPI_LAUNCH((k_elem_direct<NPE, RULE, R, PHYS, false, NT>), grid, NT, 0, s, p);
}

// (npe, rule, precision, physics) -> instantiation
template <template <int, int, typename, int> class F, typename... A>
static hipError_t dispatch(int npe, int rule, int prec, int phys, A &&...a) {
#define CPI_PHYS(NPE, RULE, R)                                              \
  switch (phys) {                                                           \
    case PHYS_COMPAT: return F<NPE, RULE, R, PHYS_COMPAT>::run(a...);      \
    case PHYS_EP: return F<NPE, RULE, R, PHYS_EP>::run(a...);              \
    case PHYS_EP_MLP: return F<NPE, RULE, R, PHYS_EP_MLP>::run(a...);      \
  }                                                                         \
  return hipErrorInvalidValue;
#define CPI_PREC(NPE, RULE)           \
  if (prec == 1) {                    \
    CPI_PHYS(NPE, RULE, float)        \
  } else {                            \
    CPI_PHYS(NPE, RULE, double)       \
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

template <int NPE, int RULE, typename R, int PHYS> struct ElemBlockF {
  static hipError_t run(const Params &p, bool aux, int nb, size_t smem, hipStream_t s) {
    if constexpr (PHYS == PHYS_EP_MLP && (NPE != 4 || !std::is_same<R, float>::value))
      return hipErrorInvalidValue;
    else
      return elem_block_t<NPE, RULE, R, PHYS>(p, aux, nb, smem, s);
  }
};
template <int NPE, int RULE, typename R, int PHYS> struct ElemBlockPrepF {
  static hipError_t run(size_t smem) {
    if constexpr (PHYS == PHYS_EP_MLP && (NPE != 4 || !std::is_same<R, float>::value)) {
      return hipErrorInvalidValue;
    } else {
      constexpr int NT = block_threads<NPE>();
      hipError_t e = lds_opt_in(k_elem_block<NPE, RULE, R, PHYS, true, NT>, smem);
      if (e != hipSuccess) return e;
      return lds_opt_in(k_elem_block<NPE, RULE, R, PHYS, false, NT>, smem);
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
                             size_t smem, hipStream_t s) {
  return dispatch<ElemBlockF>(p.npe, rule, prec, phys, p, aux, nblocks, smem, s);
}

hipError_t prepare_elem_block(int npe, int rule, int prec, int phys, size_t smem) {
  return dispatch<ElemBlockPrepF>(npe, rule, prec, phys, smem);
}

hipError_t launch_elem_direct(const Params &p, int rule, int prec, int phys, bool atomic,
                              hipStream_t s) {
  return dispatch<ElemDirectF>(p.npe, rule, prec, phys, p, atomic, s);
}

hipError_t launch_node_bnd(const Params &p, bool aux, hipStream_t s) {
  const int grid = std::max(1, (p.n_boundary + 255) / 256);
  if (aux) CPI_LAUNCH(k_node_bnd<true>, grid, 256, 0, s, p);
  CPI_LAUNCH(k_node_bnd<false>, grid, 256, 0, s, p);
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
  CPI_LAUNCH(k_energy_reduce, 1, 256, 0, s, p);
}

hipError_t launch_mass(const Params &p, int rule, int scatter, int kind, double rho, double *m,
                       int nblocks, hipStream_t s) {
  if (scatter == 0) {
    const int nt = p.npe == 4 ? 512 : 128;
    const size_t smem = 8 * (4 * (size_t)p.max_loc + (size_t)p.max_el * p.npe);
    if (smem > 65536) return hipErrorInvalidValue;  // sized by the plan to stay below
    hipError_t err;
    if (p.npe == 4) {
      switch (rule) {
        case RULE_REF4:
          hipLaunchKernelGGL((k_mass_block<4, RULE_REF4, 512>), dim3(nblocks), dim3(nt), smem, s,
                             p, kind, rho, m);
          break;
        case RULE_EXACT4:
          hipLaunchKernelGGL((k_mass_block<4, RULE_EXACT4, 512>), dim3(nblocks), dim3(nt), smem,
                             s, p, kind, rho, m);
          break;
        case RULE_CENTROID:
          hipLaunchKernelGGL((k_mass_block<4, RULE_CENTROID, 512>), dim3(nblocks), dim3(nt), smem,
                             s, p, kind, rho, m);
          break;
        default:
          hipLaunchKernelGGL((k_mass_block<4, RULE_REF5, 512>), dim3(nblocks), dim3(nt), smem, s,
                             p, kind, rho, m);
      }
    } else {
      switch (rule) {
        case RULE_REF5:
          hipLaunchKernelGGL((k_mass_block<10, RULE_REF5, 128>), dim3(nblocks), dim3(nt), smem, s,
                             p, kind, rho, m);
          break;
        case RULE_REF4:
          hipLaunchKernelGGL((k_mass_block<10, RULE_REF4, 128>), dim3(nblocks), dim3(nt), smem, s,
                             p, kind, rho, m);
          break;
        default:
          hipLaunchKernelGGL((k_mass_block<10, RULE_EXACT4, 128>), dim3(nblocks), dim3(nt), smem,
                             s, p, kind, rho, m);
      }
    }
    if ((err = hipGetLastError()) != hipSuccess) return err;
    if (p.n_boundary > 0)
      hipLaunchKernelGGL(k_mass_bnd, dim3((p.n_boundary + 255) / 256), dim3(256), 0, s, p, m);
    return hipGetLastError();
  }
  const int grid = (p.elem_count + 255) / 256;
  if (grid == 0) return hipSuccess;
  const bool at = scatter == 2;
#define CPI_MD(NPE, RULE)                                                                        \
  if (at)                                                                                        \
    hipLaunchKernelGGL((k_mass_direct<NPE, RULE, true>), dim3(grid), dim3(256), 0, s, p, kind, rho, m); \
  else                                                                                           \
    hipLaunchKernelGGL((k_mass_direct<NPE, RULE, false>), dim3(grid), dim3(256), 0, s, p, kind, rho, m);
  if (p.npe == 4) {
    if (rule == RULE_CENTROID) {
      CPI_MD(4, RULE_CENTROID)
    } else if (rule == RULE_REF4) {
      CPI_MD(4, RULE_REF4)
    } else if (rule == RULE_EXACT4) {
      CPI_MD(4, RULE_EXACT4)
    } else {
      CPI_MD(4, RULE_REF5)
    }
  } else {
    if (rule == RULE_REF4) {
      CPI_MD(10, RULE_REF4)
    } else if (rule == RULE_EXACT4) {
      CPI_MD(10, RULE_EXACT4)
    } else {
      CPI_MD(10, RULE_REF5)
    }
  }
#undef CPI_MD
  return hipGetLastError();
}

hipError_t launch_elem_dt(const Params &p, double *out, hipStream_t s) {
  const int grid = (p.ne + 255) / 256;
  if (p.npe == 4) CPI_LAUNCH(k_elem_dt<4>, grid, 256, 0, s, p, out);
  CPI_LAUNCH(k_elem_dt<10>, grid, 256, 0, s, p, out);
}

hipError_t launch_stress_elem(const Params &p, const double *const X0[3], double *out,
                              hipStream_t s) {
  const int grid = (p.ne + 255) / 256;
  if (p.npe == 4) CPI_LAUNCH(k_stress_elem<4>, grid, 256, 0, s, p, X0[0], X0[1], X0[2], out);
  CPI_LAUNCH(k_stress_elem<10>, grid, 256, 0, s, p, X0[0], X0[1], X0[2], out);
}

hipError_t launch_debug_strain(const Params &p, int rule, double alpha, int dcase, double nu,
                               double *out, hipStream_t s) {
  const int grid = (p.ne + 255) / 256;
  if (p.npe == 4) {
    switch (rule) {
      case RULE_REF4: CPI_LAUNCH((k_debug_strain<4, RULE_REF4>), grid, 256, 0, s, p, alpha, dcase, nu, out);
      case RULE_EXACT4: CPI_LAUNCH((k_debug_strain<4, RULE_EXACT4>), grid, 256, 0, s, p, alpha, dcase, nu, out);
      case RULE_CENTROID: CPI_LAUNCH((k_debug_strain<4, RULE_CENTROID>), grid, 256, 0, s, p, alpha, dcase, nu, out);
      default: CPI_LAUNCH((k_debug_strain<4, RULE_REF5>), grid, 256, 0, s, p, alpha, dcase, nu, out);
    }
  }
  switch (rule) {
    case RULE_REF4: CPI_LAUNCH((k_debug_strain<10, RULE_REF4>), grid, 256, 0, s, p, alpha, dcase, nu, out);
    case RULE_EXACT4: CPI_LAUNCH((k_debug_strain<10, RULE_EXACT4>), grid, 256, 0, s, p, alpha, dcase, nu, out);
    default: CPI_LAUNCH((k_debug_strain<10, RULE_REF5>), grid, 256, 0, s, p, alpha, dcase, nu, out);
  }
}

}  // namespace cpi
