// Context management and the C ABI (include/cppimpact_hip.h) of
// libcppimpact_hip.so: device buffers, uploads/downloads between the
// reference's host layouts and the device layout, the hipGraph-captured
// step, timing and diagnostics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cppimpact_hip.h"
#include "device.h"
#include "internal.h"
#include "kernels.h"
#include "mlp_host.h"

namespace cpi {
void set_global_error(const std::string &s);
const char *global_error();
}  // namespace cpi

using namespace cpi;

struct cpi_ctx {
  int dev = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int nn = 0, ne = 0, npe = 0, nq = 0, rule = 0, phys = 0, prec = 0, scatter = 0, mass_kind = 0;
  cpi_options opt{};
  cpi_material mat{};
  cpi_wall wall{};
  std::vector<double> X0u, V0u;  // user order [N][3]
  std::vector<int32_t> conn_u, color, slave_u;
  int ncol = 0;
  std::vector<int32_t> eperm, nperm, ninv;  // device->user elem, device->user node, user->device
  std::vector<int32_t> color_off;           // colour scatter launch ranges (device order)
  BlockPlan plan;
  std::vector<void *> allocs;
  double *d_tmp = nullptr;
  size_t tmp_len = 0;
  Params P{};
  int nblocks = 0, n_elem_rows = 0, n_node_rows = 0, n_rows = 0;
  std::vector<int> color_rows;  // first energy row of each colour launch
  size_t smem = 0;
  // captured steps: one plain step, the energy/output step, and kGraphRuns[k]
  // plain steps in one graph each
  struct GraphSet {
    hipGraphExec_t plain = nullptr, aux = nullptr;
    hipGraphExec_t run[2] = {nullptr, nullptr};
  } gs;
  bool mass_ready = false, inited = false;
  double dt = 0.0;
  int64_t step = 0;
  MlpHost mlp;  // surrogate weights and sample buffer
};

namespace {

#define HIPCHK(ctx, expr)                                                                  \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      (ctx)->err = std::string(#expr) + " failed at " __FILE__ ":" + std::to_string(__LINE__) + \
                   ": " + hipGetErrorString(e_);                                          \
      return CPI_ERR_HIP;                                                                 \
    }                                                                                     \
  } while (0)

int fail(cpi_ctx *c, int code, const std::string &msg) {
  c->err = msg;
  return code;
}

template <typename T>
int dalloc(cpi_ctx *c, T **p, size_t n) {
  void *q = nullptr;
  hipError_t e = hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T));
  if (e != hipSuccess) {
    c->err = "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed: " +
             hipGetErrorString(e);
    return CPI_ERR_ALLOC;
  }
  c->allocs.push_back(q);
  *p = (T *)q;
  if (c->opt.debug_fill_nan && c->stream)
    HIPCHK(c, hipMemsetAsync(q, 0xFF, std::max<size_t>(n, 1) * sizeof(T), c->stream));
  return CPI_OK;
}

template <typename T>
int upload(cpi_ctx *c, T *d, const std::vector<T> &h) {
  if (h.empty()) return CPI_OK;
  // on the context's stream: kernels there do not wait for the null stream,
  // and a pageable host-to-device copy may return before its DMA lands
  HIPCHK(c, hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CPI_OK;
}

int default_rule(int npe, int phys, int q) {
  if (q >= 0) return q;
  if (phys == CPI_PHYSICS_COMPAT) return CPI_QUAD_REF5;
  return npe == 4 ? CPI_QUAD_CENTROID : CPI_QUAD_EXACT4;
}

double rule_weight_sum(int rule) {
  switch (rule) {
    case CPI_QUAD_REF5: return -0.13333333333 + 0.075 + 0.075 + 0.075 + 0.075;
    case CPI_QUAD_REF4: return 0.041666667 + 0.041666667 + 0.041666667 + 0.041666667;
    case CPI_QUAD_EXACT4: return 4.0 / 24.0;
    default: return 1.0 / 6.0;
  }
}

// user AoS [N][3] <-> device nodal records in device node order (Node4
// with w = 0, or Vec3)
template <typename T>
int put_nodes(cpi_ctx *c, T *d, const double *aos) {
  std::vector<T> tmp(c->nn);
  for (int i = 0; i < c->nn; i++) {
    const double *s = aos + 3L * c->nperm[i];
    tmp[i] = T{};
    tmp[i].x = s[0];
    tmp[i].y = s[1];
    tmp[i].z = s[2];
  }
  HIPCHK(c, hipMemcpyAsync(d, tmp.data(), sizeof(T) * c->nn, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CPI_OK;
}

template <typename T>
int get_nodes(cpi_ctx *c, const T *d, double *aos) {
  std::vector<T> tmp(c->nn);
  HIPCHK(c, hipMemcpyAsync(tmp.data(), d, sizeof(T) * c->nn, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < c->nn; i++) {
    double *o = aos + 3L * c->nperm[i];
    o[0] = tmp[i].x;
    o[1] = tmp[i].y;
    o[2] = tmp[i].z;
  }
  return CPI_OK;
}

int enqueue_step(cpi_ctx *c, bool aux) {
  Params p = c->P;
  p.store_aux = aux ? 1 : 0;
  hipError_t e = hipSuccess;
  if (c->scatter == CPI_SCATTER_BLOCK) {
    p.epart_base = 0;
    e = launch_elem_block(p, c->rule, c->prec, c->phys, aux, c->nblocks, c->smem, c->stream);
    if (e == hipSuccess) {
      p.epart_base = c->n_elem_rows;
      e = launch_node_bnd(p, c->prec, aux, c->stream);
    }
  } else {
    const bool atomic = c->scatter == CPI_SCATTER_ATOMIC;
    const int nc = atomic ? 1 : c->ncol;
    for (int k = 0; k < nc && e == hipSuccess; k++) {
      p.elem_base = atomic ? 0 : c->color_off[k];
      p.elem_count = atomic ? c->ne : c->color_off[k + 1] - c->color_off[k];
      p.epart_base = c->color_rows[k];
      e = launch_elem_direct(p, c->rule, c->prec, c->phys, atomic, c->stream);
    }
    if (e == hipSuccess) {
      p.epart_base = c->n_elem_rows;
      e = launch_node_all(p, aux, c->stream);
    }
  }
  if (e == hipSuccess && aux) e = launch_energy_reduce(p, c->stream);
  if (e != hipSuccess) return fail(c, CPI_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return CPI_OK;
}

// plain steps captured back to back in GraphSet::run: one graph launch, and
// no gap for the launch between the steps' kernels, per kGraphRuns[k]
// steps (measured at c2: ≈ 4 µs less per step than one graph per step)
constexpr int kGraphRuns[2] = {32, 4};

int capture(cpi_ctx *c, bool aux, hipGraphExec_t *out, int nrep = 1) {
  hipGraph_t g = nullptr;
  HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rc = CPI_OK;
  for (int i = 0; i < nrep && rc == CPI_OK; i++) rc = enqueue_step(c, aux);
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail(c, CPI_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
  e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return fail(c, CPI_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
  return CPI_OK;
}

void drop_graphs(cpi_ctx *c) {
  auto &G = c->gs;
  for (hipGraphExec_t *g : {&G.plain, &G.aux, &G.run[0], &G.run[1]}) {
    if (*g) (void)hipGraphExecDestroy(*g);
    *g = nullptr;
  }
}

// the step graphs (see cpi_ctx::gs)
int capture_graphs(cpi_ctx *c) {
  auto &G = c->gs;
  int rc;
  if ((rc = capture(c, false, &G.plain)) || (rc = capture(c, true, &G.aux)) ||
      (rc = capture(c, false, &G.run[0], kGraphRuns[0])) ||
      (rc = capture(c, false, &G.run[1], kGraphRuns[1])))
    return rc;
  return CPI_OK;
}

int ensure_tmp(cpi_ctx *c, size_t n) {
  if (c->tmp_len >= n) return CPI_OK;
  double *p = nullptr;
  int rc = dalloc(c, &p, n);
  if (rc) return rc;
  c->d_tmp = p;
  c->tmp_len = n;
  return CPI_OK;
}

int compute_mass(cpi_ctx *c) {
  Params p = c->P;
  HIPCHK(c, hipMemsetAsync((void *)p.m, 0, sizeof(double) * c->nn, c->stream));
  const double rho = c->mat.rho;
  hipError_t e = hipSuccess;
  if (c->scatter == CPI_SCATTER_BLOCK) {
    e = launch_mass(p, c->rule, 0, c->mass_kind, rho, c->nblocks, c->stream);
  } else if (c->scatter == CPI_SCATTER_ATOMIC) {
    p.elem_base = 0;
    p.elem_count = c->ne;
    e = launch_mass(p, c->rule, 2, c->mass_kind, rho, 0, c->stream);
  } else {
    for (int k = 0; k < c->ncol && e == hipSuccess; k++) {
      p.elem_base = c->color_off[k];
      p.elem_count = c->color_off[k + 1] - c->color_off[k];
      e = launch_mass(p, c->rule, 1, c->mass_kind, rho, 0, c->stream);
    }
  }
  if (e != hipSuccess) return fail(c, CPI_ERR_HIP, std::string("mass kernel: ") + hipGetErrorString(e));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->mass_ready = true;
  return CPI_OK;
}

// x = X0, v = v0, forces/aux/state zeroed (the mass, once assembled, stays)
int reset_state(cpi_ctx *c) {
  int rc = put_nodes(c, c->P.pos, c->X0u.data());
  if (!rc) rc = put_nodes(c, c->P.vv, c->V0u.data());
  if (!rc) rc = put_nodes(c, c->P.xprev, c->X0u.data());
  if (rc) return rc;
  for (int k = 0; k < 3; k++)
    HIPCHK(c, hipMemsetAsync(c->P.f[k], 0, sizeof(double) * c->nn, c->stream));
  HIPCHK(c, hipMemsetAsync(c->P.acc, 0, sizeof(Node4) * c->nn, c->stream));
  HIPCHK(c, hipMemsetAsync(c->P.du, 0, sizeof(Node4) * c->nn, c->stream));
  const size_t rs = c->prec == CPI_PREC_F32 ? 4 : 8;
  if (c->phys != CPI_PHYSICS_COMPAT) {
    HIPCHK(c, hipMemsetAsync(c->P.sig, 0, rs * 6 * c->nq * (size_t)c->ne, c->stream));
    HIPCHK(c, hipMemsetAsync(c->P.epbar, 0, rs * c->nq * (size_t)c->ne, c->stream));
    if (c->P.yielded)
      HIPCHK(c, hipMemsetAsync(c->P.yielded, 0, 8 * (size_t)c->P.ywords * c->nblocks, c->stream));
    if (c->P.mlp_ctr) HIPCHK(c, hipMemsetAsync(c->P.mlp_ctr, 0, sizeof(unsigned long long), c->stream));
  }
  HIPCHK(c, hipMemsetAsync(c->P.epart, 0, sizeof(double) * EP_N * c->n_rows, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return CPI_OK;
}

// stable dt of the current device positions: element limit and the wall
// spring limit 2 sqrt(m_min / k)
int stable_dt(cpi_ctx *c, double *dte, double *dtc) {
  const int nb = (c->ne + 255) / 256;
  int rc = ensure_tmp(c, std::max(nb, c->nn));
  if (rc) return rc;
  hipError_t e = launch_elem_dt(c->P, c->d_tmp, c->stream);
  if (e != hipSuccess) return fail(c, CPI_ERR_HIP, hipGetErrorString(e));
  std::vector<double> h(std::max(nb, c->nn));
  HIPCHK(c, hipMemcpyAsync(h.data(), c->d_tmp, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double m1 = 1e300;
  for (int i = 0; i < nb; i++) m1 = std::min(m1, h[i]);
  *dte = m1;
  HIPCHK(c, hipMemcpyAsync(h.data(), c->P.m, sizeof(double) * c->nn, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double mmin = 1e300;
  for (int i = 0; i < c->nn; i++)
    if (h[i] > 0.0) mmin = std::min(mmin, h[i]);
  *dtc = (c->wall.enabled && c->wall.stiffness > 0.0) ? 2.0 * std::sqrt(mmin / c->wall.stiffness)
                                                      : INFINITY;
  return CPI_OK;
}

int check_flags(cpi_ctx *c) {
  Ctl h;
  HIPCHK(c, hipMemcpyAsync(&h, c->P.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (h.flags & 2u) return fail(c, CPI_ERR_NUMERIC, "inverted or degenerate element (detJ <= 0) during the step");
  if (h.flags & 1u) return fail(c, CPI_ERR_NUMERIC, "non-finite value in the energy terms (NaN/Inf state)");
  if (h.flags & 4u)
    return fail(c, CPI_ERR_NUMERIC,
                "time step " + std::to_string(h.dt) + " above the stable limit " +
                    std::to_string(std::min(h.row[7], h.dt_contact)) +
                    " of the configuration (element CFL / wall spring)");
  return CPI_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" void cpi_default_options(cpi_options *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->physics = CPI_PHYSICS_COMPAT;
  o->precision = CPI_PREC_F64;
  o->scatter = CPI_SCATTER_BLOCK;
  o->quadrature = CPI_QUAD_DEFAULT;
  o->mass = CPI_MASS_DEFAULT;
  o->stress_update = CPI_RETURN_EXACT;
  o->use_graph = 1;
  o->track_energy = 1;
  o->contact_slave_only = 0;
  o->fixed_bc = 0;
  o->debug_fill_nan = 0;
  o->stream_state = 0;
  o->block_elements = 0;
  o->device = 0;
  o->gravity[0] = 0.0;
  o->gravity[1] = 0.0;
  o->gravity[2] = -9.81;
  o->mlp_tolerance = 1e-3;
}

extern "C" const char *cpi_last_error(const cpi_ctx *c) {
  return c ? c->err.c_str() : cpi::global_error();
}

extern "C" void cpi_destroy(cpi_ctx *c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graphs(c);
  for (void *p : c->allocs) (void)hipFree(p);
  mlp_free(c->mlp);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

// The block kernels that read each corner as a 16-bit word, l | k << 9
// (local node, its contribution rank): the per-point tet10 kernel always,
// the tet4 kernels when the plan's blocks fit the words.
static bool corner_words(const cpi_ctx *c) {
  return c->scatter == CPI_SCATTER_BLOCK &&
         ((c->npe == 10 && c->nq == 4) || (c->npe == 4 && c->plan.corner_fits()));
}

static int create_impl(cpi_ctx *c, const cpi_mesh_desc *m, const cpi_material *mat,
                       const cpi_wall *wall, const cpi_options *opt) {
  c->opt = *opt;
  c->mat = *mat;
  if (wall) c->wall = *wall;
  c->nn = m->num_nodes;
  c->ne = m->num_elements;
  c->npe = m->nodes_per_element;
  const int npe = c->npe;
  if (c->nn <= 0 || c->ne <= 0 || (npe != 4 && npe != 10) || !m->xloc || !m->element_nodes)
    return fail(c, CPI_ERR_INVALID, "cpi_create: empty mesh or nodes_per_element not 4/10");
  if (opt->physics < 0 || opt->physics > 1 || opt->precision < 0 || opt->precision > 1 ||
      opt->scatter < 0 || opt->scatter > 2 || opt->stress_update < 0 || opt->stress_update > 1 ||
      opt->stream_state < 0 || opt->stream_state > 2)
    return fail(c, CPI_ERR_INVALID, "cpi_create: option out of range");
  if (!(mat->E > 0) || !(mat->rho > 0) || !(mat->nu > -1.0 && mat->nu < 0.5))
    return fail(c, CPI_ERR_INVALID, "cpi_create: material needs E > 0, rho > 0, -1 < nu < 0.5");
  if (wall && wall->enabled && (wall->dim < 0 || wall->dim > 2 || (wall->normal != 1 && wall->normal != -1)))
    return fail(c, CPI_ERR_INVALID, "cpi_create: wall dim must be 0..2 and normal +-1");
  c->phys = opt->physics;
  if (c->phys == CPI_PHYSICS_EP && opt->stress_update == CPI_RETURN_MLP) {
    if (npe != 4 || opt->precision != CPI_PREC_F32 || opt->scatter != CPI_SCATTER_BLOCK)
      return fail(c, CPI_ERR_INVALID, "the MLP surrogate runs on tet4, FP32, block scatter");
    c->phys = 2;
  }
  c->prec = opt->precision;
  c->scatter = opt->scatter;
  c->rule = default_rule(npe, opt->physics, opt->quadrature);
  if (c->rule < 0 || c->rule > 3) return fail(c, CPI_ERR_INVALID, "cpi_create: bad quadrature");
  if (npe == 10 && c->rule == CPI_QUAD_CENTROID)
    return fail(c, CPI_ERR_INVALID, "tet10 needs a 4- or 5-point rule");
  c->mass_kind = opt->mass >= 0 ? opt->mass
                                 : (opt->physics == CPI_PHYSICS_COMPAT ? CPI_MASS_ROWSUM : CPI_MASS_HRZ);
  c->nq = (npe == 4) ? 1 : (c->rule == CPI_QUAD_REF5 ? 5 : 4);
  // host copies and validation
  c->X0u.assign(m->xloc, m->xloc + 3L * c->nn);
  c->V0u.assign(3L * c->nn, 0.0);
  c->conn_u.assign(m->element_nodes, m->element_nodes + (long)npe * c->ne);
  for (int32_t v : c->conn_u)
    if (v < 0 || v >= c->nn) return fail(c, CPI_ERR_MESH, "element_nodes references a node out of range");
  for (double v : c->X0u)
    if (!std::isfinite(v)) return fail(c, CPI_ERR_MESH, "non-finite node coordinate");
  for (int e = 0; e < c->ne; e++) {
    const int32_t *ce = &c->conn_u[(long)npe * e];
    const int o = npe == 4 ? 0 : 2, i1 = npe == 4 ? 1 : 0, i2 = npe == 4 ? 2 : 1;
    double a[3], b[3], d[3];
    for (int k = 0; k < 3; k++) {
      a[k] = c->X0u[3L * ce[i1] + k] - c->X0u[3L * ce[o] + k];
      b[k] = c->X0u[3L * ce[i2] + k] - c->X0u[3L * ce[o] + k];
      d[k] = c->X0u[3L * ce[3] + k] - c->X0u[3L * ce[o] + k];
    }
    double v6 = a[0] * (b[1] * d[2] - b[2] * d[1]) - a[1] * (b[0] * d[2] - b[2] * d[0]) +
                a[2] * (b[0] * d[1] - b[1] * d[0]);
    if (!(v6 > 0.0))
      return fail(c, CPI_ERR_MESH, "element " + std::to_string(e) + " is inverted or degenerate");
  }
  // colouring (always: reported by cpi_get_coloring)
  c->color.assign(c->ne, 0);
  c->ncol = greedy_coloring(c->ne, npe, c->conn_u.data(), c->nn, c->color.data());
  // device
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= opt->device || opt->device < 0)
    return fail(c, CPI_ERR_NO_DEVICE, "no HIP device " + std::to_string(opt->device));
  HIPCHK(c, hipSetDevice(opt->device));
  hipDeviceProp_t prop;
  HIPCHK(c, hipGetDeviceProperties(&prop, opt->device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return fail(c, CPI_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName +
                                          ", this library is built for gfx950 only");
  c->dev = opt->device;
  HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  // device orders
  if (c->scatter == CPI_SCATTER_BLOCK) {
    const int nt = elem_block_threads(npe, c->rule);
    const int dflt = elem_block_default_elements(npe);
    int me = opt->block_elements > 0 ? opt->block_elements : dflt;
    // the tet4 kernel gives each thread one element and at most one node,
    // the tet10 kernel of the 4-point rules each thread one point
    if (npe == 4 && me > nt)
      return fail(c, CPI_ERR_INVALID, "block_elements above " + std::to_string(nt) + " for tet4");
    if (npe == 10 && c->nq == 4 && 4 * me > nt)
      return fail(c, CPI_ERR_INVALID, "block_elements above " + std::to_string(nt / 4) +
                                          " for tet10 on a 4-point rule");
    // LDS budget per workgroup: three (FP64) or four (FP32) tet4 workgroups,
    // two tet10 ones per 160 KiB CU, never above the device's limit; node
    // records take 56 bytes each (slot offsets budgeted for 64 contributions
    // per node)
    const long share = npe == 4 ? (c->prec == CPI_PREC_F32 ? 40960 : 54272) : 81920;
    const long budget = std::min<long>((long)prop.sharedMemPerBlock, share);
    const long fixed = (long)elem_block_smem(npe, c->prec, c->phys, 0, me, 64);
    long mnl = (budget - fixed - 16) / 56;
    if (mnl < 4L * npe)
      return fail(c, CPI_ERR_INVALID, "block_elements " + std::to_string(me) +
                                          " does not fit the LDS budget");
    // the tet4 and the per-point tet10 kernels give each thread at most one
    // node; the five-point tet10 kernel loops over its nodes
    int mn = (int)std::min<long>(mnl, (npe == 4 || c->nq == 4) ? nt : 1024);
    std::string e;
    if (!build_block_plan(c->ne, npe, c->conn_u.data(), c->nn, c->X0u.data(), me, mn, c->plan, e))
      return fail(c, CPI_ERR_INVALID, "partition: " + e);
    if (c->plan.max_local_nodes > mn)
      return fail(c, CPI_ERR_INVALID, "partition: a block holds more nodes than its kernel");
    // the per-point tet10 kernel reads each corner as a 16-bit word (local
    // node < 512, contribution rank < 128), which its limits of 128
    // elements and 512 nodes per block guarantee
    if (npe == 10 && c->nq == 4 && !c->plan.corner_fits())
      return fail(c, CPI_ERR_INVALID, "partition: a tet10 block exceeds its corner words");
    c->eperm = c->plan.elem_perm;
    c->nperm = c->plan.node_perm;
    c->ninv = c->plan.node_inv;
    c->nblocks = c->plan.nblocks;
  } else {
    c->nperm.resize(c->nn);
    c->ninv.resize(c->nn);
    for (int i = 0; i < c->nn; i++) c->nperm[i] = c->ninv[i] = i;
    if (c->scatter == CPI_SCATTER_COLOR) {
      color_order(c->ne, c->color.data(), c->ncol, c->eperm, c->color_off);
    } else {
      c->eperm.resize(c->ne);
      for (int e = 0; e < c->ne; e++) c->eperm[e] = e;
    }
  }
  Params &P = c->P;
  P.nn = c->nn;
  P.ne = c->ne;
  P.npe = npe;
  P.nq = c->nq;
  const double E = mat->E, nu = mat->nu;
  const double k = E / ((1 + nu) * (1 - 2 * nu));
  P.c11 = (1 - nu) * k;
  P.c12 = nu * k;
  P.c44 = ((1 - 2 * nu) / 2) * k;
  P.mu = E / (2.0 * (1.0 + nu));
  P.sy0 = mat->Y0;
  P.H = mat->H;
  P.wave_speed = std::sqrt(P.c11 / mat->rho);
  P.wall_on = (wall && wall->enabled) ? 1 : 0;
  P.wall_dim = wall ? wall->dim : 2;
  P.wall_normal = wall ? wall->normal : 1;
  P.wall_loc = wall ? wall->location : 0.0;
  P.wall_k = wall ? wall->stiffness : 0.0;
  P.wall_nk = wall ? wall->normal * wall->stiffness : 0.0;
  for (int i = 0; i < 3; i++) P.grav[i] = opt->gravity[i];
  P.wsum = rule_weight_sum(c->rule);
  int rc = CPI_OK;
  const int nn = c->nn, ne = c->ne;
  for (int kk = 0; kk < 3 && !rc; kk++) {
    rc = dalloc(c, &P.f[kk], nn);
  }
  if (!rc) rc = dalloc(c, &P.pos, nn);
  if (!rc) rc = dalloc(c, &P.vv, nn);
  if (!rc) rc = dalloc(c, &P.acc, nn);
  if (!rc) rc = dalloc(c, &P.du, nn);
  if (!rc) rc = dalloc(c, &P.xprev, nn);
  if (!rc) rc = dalloc(c, &P.X0, nn);
  double *dm = nullptr;
  if (!rc) rc = dalloc(c, &dm, nn);
  P.m = dm;
  if (rc) return rc;
  // connectivity [npe][ne] in device element and node order
  {
    std::vector<int32_t> cd((size_t)npe * ne);
    for (int e = 0; e < ne; e++)
      for (int j = 0; j < npe; j++)
        cd[(size_t)j * ne + e] = c->ninv[c->conn_u[(size_t)npe * c->eperm[e] + j]];
    int32_t *d = nullptr;
    if ((rc = dalloc(c, &d, cd.size())) || (rc = upload(c, d, cd))) return rc;
    P.conn = d;
  }
  // per-node flags in device node order from a list of user node ids
  auto node_mask = [&](int n, const int32_t *ids, const char *what, const uint8_t **out) {
    std::vector<uint8_t> mask(nn, 0);
    for (int i = 0; i < n; i++) {
      const int s = ids ? ids[i] : -1;
      if (s < 0 || s >= nn) return fail(c, CPI_ERR_MESH, std::string(what) + " node out of range");
      mask[c->ninv[s]] = 1;
    }
    uint8_t *d = nullptr;
    int r = dalloc(c, &d, mask.size());
    if (!r) r = upload(c, d, mask);
    if (!r) *out = d;
    return r;
  };
  if (opt->contact_slave_only &&
      (rc = node_mask(m->num_slave_nodes, m->slave_nodes, "slave", &P.contact_mask)))
    return rc;
  if (opt->fixed_bc && (rc = node_mask(m->num_fixed_nodes, m->fixed_nodes, "fixed", &P.fixed_mask)))
    return rc;
  if (c->phys != CPI_PHYSICS_COMPAT) {
    const size_t rs = c->prec == CPI_PREC_F32 ? 4 : 8;
    char *s = nullptr, *q = nullptr;
    if ((rc = dalloc(c, &s, rs * 6 * c->nq * (size_t)ne)) || (rc = dalloc(c, &q, rs * c->nq * (size_t)ne)))
      return rc;
    P.sig = s;
    P.epbar = q;
    // when a step's working set (state, node records, partial forces,
    // connectivity) exceeds 192 MB, three quarters of the 256 MB Infinity
    // Cache, the block kernels stream the state past the caches, which then
    // keep the node records and partials that other blocks and the node
    // pass read again: the c3 cylinder (285 MB in FP32, 440 MB in FP64) and
    // the 128^3 tet4 cube (0.6-1 GB) step 4-6% faster; the c2 cube (89 and
    // 138 MB) stays cached, which streaming would make 6-9% slower
    const size_t slots = c->scatter == CPI_SCATTER_BLOCK ? (size_t)c->plan.n_slots : 0;
    const size_t conn = corner_words(c) ? 2 * (size_t)npe : 4 * (size_t)npe;
    const size_t ws = rs * 7 * c->nq * (size_t)ne + 56 * (size_t)c->nn +
                      slots * (c->prec == CPI_PREC_F32 ? 12 : 24) + conn * ne;
    P.stream_state = opt->stream_state == 1   ? 1
                     : opt->stream_state == 2 ? 0
                                              : (ws > ((size_t)192 << 20));
    // the tet4 kernels of plans beyond the corner words keep it cached
    if (npe == 4 && c->scatter == CPI_SCATTER_BLOCK && !corner_words(c)) P.stream_state = 0;
  }
  if (c->scatter == CPI_SCATTER_BLOCK) {
    BlockPlan &B = c->plan;
    if (npe == 4 && B.max_local_nodes > elem_block_threads(4, c->rule))
      return fail(c, CPI_ERR_INVALID, "partition: a tet4 block has more nodes than threads");
    if (npe == 10 && c->nq == 4 && 4 * B.max_block_elems > elem_block_threads(10, c->rule))
      return fail(c, CPI_ERR_INVALID, "partition: a tet10 block has more points than threads");
    int32_t *meta, *bn, *bs, *sp, *dn, *ds;
    uint16_t *cl, *ci, *cw = nullptr, *dg, *dof;
    Vec3 *part;
    const bool words = corner_words(c);
    if ((rc = dalloc(c, &meta, B.meta.size())) || (rc = upload(c, meta, B.meta)) ||
        (rc = dalloc(c, &cl, B.conn_local.size())) || (rc = upload(c, cl, B.conn_local)) ||
        (rc = dalloc(c, &ci, B.csr_pos.size())) || (rc = upload(c, ci, B.csr_pos)) ||
        (words && ((rc = dalloc(c, &cw, B.corner.size())) || (rc = upload(c, cw, B.corner)))) ||
        (rc = dalloc(c, &bn, B.bnd_nodes.size())) || (rc = upload(c, bn, B.bnd_nodes)) ||
        (rc = dalloc(c, &bs, B.bnd_slot.size())) || (rc = upload(c, bs, B.bnd_slot)) ||
        (rc = dalloc(c, &sp, B.slot_ptr.size())) || (rc = upload(c, sp, B.slot_ptr)) ||
        (rc = dalloc(c, &dn, B.nid_pad.size())) || (rc = upload(c, dn, B.nid_pad)) ||
        (rc = dalloc(c, &ds, B.slot_pad.size())) || (rc = upload(c, ds, B.slot_pad)) ||
        (rc = dalloc(c, &dg, B.deg_pad.size())) || (rc = upload(c, dg, B.deg_pad)) ||
        (rc = dalloc(c, &dof, B.off_pad.size())) || (rc = upload(c, dof, B.off_pad)) ||
        (rc = dalloc(c, &part, (size_t)std::max(B.n_slots, 1))))
      return rc;
    P.meta = meta;
    P.conn_local = cl;
    P.csr_pos = ci;
    P.corner = cw;
    P.bnd_nodes = bn;
    P.bnd_slot = bs;
    P.slot_ptr = sp;
    P.nid_pad = dn;
    P.slot_pad = ds;
    P.deg_pad = dg;
    P.off_pad = dof;
    P.ml = B.max_local_nodes;
    P.kmax = B.kmax;
    P.partial = part;
    P.n_interior = B.n_interior;
    P.n_boundary = B.n_boundary;
    P.max_loc = B.max_local_nodes;
    P.max_el = B.max_block_elems;
    // the yielded bits, one per stress point: tet4 and the per-point tet10
    // kernel, exact return (the surrogate kernel reads ep always and keeps
    // no bits)
    if ((npe == 4 || c->nq == 4) && c->phys == CPI_PHYSICS_EP) {
      unsigned long long *yw = nullptr;
      P.ywords = (c->nq * B.max_block_elems + 63) / 64;
      if ((rc = dalloc(c, &yw, (size_t)P.ywords * B.nblocks))) return rc;
      P.yielded = yw;
    }
    // surrogate mode: the step's counter, the deferred-block list and the
    // network queue (one row per element at most, plus a tile of padding
    // the last tile reads)
    if (npe == 4 && c->phys == 2) {
      unsigned long long *ctr = nullptr;
      int2 *lst = nullptr;
      uint16_t *rows = nullptr;
      const size_t nrows = (size_t)ne + 16;
      if ((rc = dalloc(c, &ctr, 1)) || (rc = dalloc(c, &lst, (size_t)B.nblocks)) ||
          (rc = dalloc(c, &rows, nrows * 16)))
        return rc;
      HIPCHK(c, hipMemsetAsync(rows, 0, nrows * 32, c->stream));
      P.mlp_ctr = ctr;
      P.mlp_list = lst;
      P.mlp_rows = rows;
    }
    P.nblocks = B.nblocks;
    c->smem = elem_block_smem(npe, c->prec, c->phys, P.ml, P.max_el, P.kmax);
    if (c->smem > (size_t)prop.sharedMemPerBlock)
      return fail(c, CPI_ERR_INVALID, "element block needs " + std::to_string(c->smem) +
                                          " bytes of LDS; lower block_elements");
    hipError_t pe = prepare_elem_block(npe, c->rule, c->prec, c->phys, c->smem);
    if (pe != hipSuccess)
      return fail(c, CPI_ERR_HIP, std::string("LDS opt-in for the element kernel: ") +
                                      hipGetErrorString(pe));
    // one energy row per wave
    c->n_elem_rows = c->nblocks * (elem_block_threads(npe, c->rule) / 64);
    c->n_node_rows = 4 * std::max(1, (B.n_boundary + 255) / 256);
  } else {
    const int nt = elem_direct_threads(npe);
    c->color_rows.clear();
    int rows = 0;
    if (c->scatter == CPI_SCATTER_COLOR) {
      for (int kk = 0; kk < c->ncol; kk++) {
        c->color_rows.push_back(rows);
        rows += (nt / 64) * ((c->color_off[kk + 1] - c->color_off[kk] + nt - 1) / nt);
      }
    } else {
      c->color_rows.push_back(0);
      rows = (nt / 64) * ((ne + nt - 1) / nt);
    }
    c->n_elem_rows = rows;
    c->n_node_rows = 4 * ((nn + 255) / 256);
  }
  c->n_rows = std::max(c->n_elem_rows + c->n_node_rows, 4 * ((nn + 255) / 256));
  double *ep = nullptr, *hist = nullptr, *ered = nullptr;
  Ctl *ctl = nullptr;
  if ((rc = dalloc(c, &ep, (size_t)EP_N * c->n_rows)) || (rc = dalloc(c, &ctl, 1)) ||
      (rc = dalloc(c, &hist, 8 * 4096)) ||
      (rc = dalloc(c, &ered, (size_t)EP_N * energy_reduce_blocks())))
    return rc;
  P.ered = ered;
  P.epart = ep;
  P.n_epart = c->n_rows;
  P.ctl = ctl;
  P.hist = hist;
  P.hist_cap = 4096;
  P.mlp_tol = opt->mlp_tolerance;
  P.sample_keep = 1ull << 32;  // the recorder keeps every yielding point
  if (c->phys == 2) {
    if ((rc = mlp_alloc(c->mlp, c->nblocks, c->stream, c->err))) return rc;
    P.mlp = c->mlp.d_params;
    P.mlp_stats = c->mlp.d_stats;
  }
  // reference positions on the device (stress output)
  if ((rc = put_nodes(c, c->P.X0, c->X0u.data()))) return rc;
  return reset_state(c);
}

extern "C" int cpi_create(const cpi_mesh_desc *m, const cpi_material *mat, const cpi_wall *wall,
                          const cpi_options *opt, cpi_ctx **out) {
  if (!out) {
    set_global_error("cpi_create: out is NULL");
    return CPI_ERR_INVALID;
  }
  *out = nullptr;
  if (!m || !mat) {
    set_global_error("cpi_create: mesh and material are required");
    return CPI_ERR_INVALID;
  }
  cpi_options def;
  cpi_default_options(&def);
  cpi_ctx *c = new (std::nothrow) cpi_ctx;
  if (!c) {
    set_global_error("cpi_create: out of host memory");
    return CPI_ERR_ALLOC;
  }
  int rc;
  try {
    rc = create_impl(c, m, mat, wall, opt ? opt : &def);
  } catch (const std::bad_alloc &) {
    c->err = "cpi_create: out of host memory";
    rc = CPI_ERR_ALLOC;
  }
  if (rc) {
    set_global_error(c->err);
    cpi_destroy(c);
    return rc;
  }
  *out = c;
  return CPI_OK;
}

extern "C" int cpi_set_state(cpi_ctx *c, const double *x0, const double *v0) {
  if (!c) return CPI_ERR_INVALID;
  if (x0) {
    for (long i = 0; i < 3L * c->nn; i++)
      if (!std::isfinite(x0[i])) return fail(c, CPI_ERR_INVALID, "non-finite position");
    c->X0u.assign(x0, x0 + 3L * c->nn);
    c->mass_ready = false;
  }
  if (v0) {
    for (long i = 0; i < 3L * c->nn; i++)
      if (!std::isfinite(v0[i])) return fail(c, CPI_ERR_INVALID, "non-finite velocity");
    c->V0u.assign(v0, v0 + 3L * c->nn);
  }
  c->inited = false;
  drop_graphs(c);
  int rc = put_nodes(c, c->P.X0, c->X0u.data());
  if (rc) return rc;
  return reset_state(c);
}

extern "C" int cpi_stable_dt(cpi_ctx *c, double *dte, double *dtc) {
  if (!c || !dte || !dtc) return CPI_ERR_INVALID;
  if (!c->inited) {
    int rc = reset_state(c);
    if (!rc && !c->mass_ready) rc = compute_mass(c);
    if (rc) return rc;
  }
  return stable_dt(c, dte, dtc);
}

// ---------------------------------------------------------------------------
// cpi_stable_dt without allocation: the scratch of its reductions is a
// file-scope device array, one copy per GPU, held until the process ends.
// It takes the per-workgroup element minima of k_elem_dt (one per 256
// elements) and, in its last kNoallocOut words, the two results of
// k_dt_finish.  Calls that overlap on one GPU share it and race.
// ---------------------------------------------------------------------------
constexpr int kNoallocPartials = 65536;  // 16,777,216 elements at most
constexpr int kNoallocOut = 2;
static __device__ __align__(256) double scratch[kNoallocPartials + kNoallocOut];

// one workgroup: the minimum of the element partials and the smallest
// positive nodal mass (the wall spring's limit needs it)
__global__ void __launch_bounds__(256) k_dt_finish(const double *part, int nb, const double *m,
                                                   int nn, double *out) {
  __shared__ double r0[256], r1[256];
  double a = 1e300, b = 1e300;
  for (int i = threadIdx.x; i < nb; i += 256) a = fmin(a, part[i]);
  for (int i = threadIdx.x; i < nn; i += 256)
    if (m[i] > 0.0) b = fmin(b, m[i]);
  r0[threadIdx.x] = a;
  r1[threadIdx.x] = b;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      r0[threadIdx.x] = fmin(r0[threadIdx.x], r0[threadIdx.x + h]);
      r1[threadIdx.x] = fmin(r1[threadIdx.x], r1[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = r0[0];
    out[1] = r1[0];
  }
}

namespace {

constexpr int kNoallocDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kNoallocDevices] = {};

// this device's copy of scratch, looked up once with the device current
int scratch_of(cpi_ctx *c, double **out) {
  if (c->dev < 0 || c->dev >= kNoallocDevices)
    return fail(c, CPI_ERR_INVALID, "cpi_stable_dt_noalloc: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[c->dev]) {
    HIPCHK(c, hipSetDevice(c->dev));
    void *ptr = nullptr;
    HIPCHK(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch[c->dev] = static_cast<double *>(ptr);
  }
  *out = g_scratch[c->dev];
  return CPI_OK;
}

}  // namespace

extern "C" int cpi_stable_dt_noalloc(cpi_ctx *c, double *dte, double *dtc) {
  if (!c || !dte || !dtc) return CPI_ERR_INVALID;
  if (!c->inited)
    return fail(c, CPI_ERR_STATE, "cpi_stable_dt_noalloc: needs cpi_init (mass assembled)");
  const int nb = (c->ne + 255) / 256;
  if (nb > kNoallocPartials)
    return fail(c, CPI_ERR_INVALID,
                "cpi_stable_dt_noalloc: at most " + std::to_string(256L * kNoallocPartials) +
                    " elements (the mesh has " + std::to_string(c->ne) + ")");
  double *sp = nullptr;
  int rc = scratch_of(c, &sp);
  if (rc) return rc;
  HIPCHK(c, launch_elem_dt(c->P, sp, c->stream));
  hipLaunchKernelGGL(k_dt_finish, dim3(1), dim3(256), 0, c->stream, (const double *)sp, nb,
                     (const double *)c->P.m, c->nn, sp + kNoallocPartials);
  HIPCHK(c, hipGetLastError());
  double r[kNoallocOut];
  HIPCHK(c, hipMemcpyAsync(r, sp + kNoallocPartials, sizeof(r), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *dte = r[0];
  *dtc = (c->wall.enabled && c->wall.stiffness > 0.0) ? 2.0 * std::sqrt(r[1] / c->wall.stiffness)
                                                      : INFINITY;
  return CPI_OK;
}

extern "C" int cpi_init(cpi_ctx *c, double dt) {
  if (!c) return CPI_ERR_INVALID;
  if (!std::isfinite(dt)) return fail(c, CPI_ERR_INVALID, "cpi_init: dt must be finite");
  if (c->phys == 2 && !c->mlp.ready)
    return fail(c, CPI_ERR_STATE, "cpi_init: CPI_RETURN_MLP needs cpi_set_surrogate first");
  drop_graphs(c);
  c->inited = false;
  int rc = reset_state(c);
  if (!rc) rc = compute_mass(c);
  if (rc) return rc;
  double dtc = INFINITY;
  if (dt <= 0.0) {
    double dte;
    if ((rc = stable_dt(c, &dte, &dtc))) return rc;
    const double ds = 1.0 / std::sqrt(1.0 / (dte * dte) + 1.0 / (dtc * dtc));
    dt = (dt == 0.0 ? 0.5 : -dt) * ds;
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(c, CPI_ERR_NUMERIC, "stable dt is not positive");
  } else if (c->phys != CPI_PHYSICS_COMPAT) {
    double dte;
    if ((rc = stable_dt(c, &dte, &dtc))) return rc;
  }
  c->dt = dt;
  c->step = 0;
  Ctl h{};
  h.dt = dt;
  h.dt_contact = dtc;
  HIPCHK(c, hipMemcpyAsync(c->P.ctl, &h, sizeof(Ctl), hipMemcpyHostToDevice, c->stream));
  hipError_t e = launch_node_init(c->P, c->stream);
  if (e == hipSuccess) e = launch_energy_reduce(c->P, c->stream);
  if (e != hipSuccess) return fail(c, CPI_ERR_HIP, std::string("init kernels: ") + hipGetErrorString(e));
  HIPCHK(c, hipMemsetAsync(c->P.epart, 0, sizeof(double) * EP_N * c->n_rows, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  if (c->opt.use_graph && (rc = capture_graphs(c))) return rc;
  c->inited = true;
  return CPI_OK;
}

extern "C" int cpi_step(cpi_ctx *c, int32_t nsteps) {
  if (!c) return CPI_ERR_INVALID;
  if (!c->inited) return fail(c, CPI_ERR_STATE, "cpi_step before cpi_init");
  if (nsteps < 0) return fail(c, CPI_ERR_INVALID, "nsteps < 0");
  for (int i = 0; i < nsteps;) {
    const bool aux = (i == nsteps - 1);
    const auto &G = c->gs;
    if (G.plain) {
      int k = 0;
      while (k < 2 && !(G.run[k] && nsteps - 1 - i >= kGraphRuns[k])) k++;
      if (!aux && k < 2) {
        HIPCHK(c, hipGraphLaunch(G.run[k], c->stream));
        i += kGraphRuns[k];
        continue;
      }
      HIPCHK(c, hipGraphLaunch(aux ? G.aux : G.plain, c->stream));
    } else {
      int rc = enqueue_step(c, aux);
      if (rc) return rc;
    }
    i++;
  }
  c->step += nsteps;
  return CPI_OK;
}

extern "C" int cpi_synchronize(cpi_ctx *c) {
  if (!c) return CPI_ERR_INVALID;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return check_flags(c);
}

extern "C" int cpi_get_state(cpi_ctx *c, double *x, double *v, double *a, double *m, double *vi) {
  if (!c) return CPI_ERR_INVALID;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int rc = CPI_OK;
  if (x && (rc = get_nodes(c, c->P.pos, x))) return rc;
  std::vector<double> hv, ha;
  if (v || vi) {
    hv.resize(3L * c->nn);
    if ((rc = get_nodes(c, c->P.vv, hv.data()))) return rc;
    if (v) std::memcpy(v, hv.data(), sizeof(double) * hv.size());
  }
  if (a || vi) {
    ha.resize(3L * c->nn);
    if ((rc = get_nodes(c, c->P.acc, ha.data()))) return rc;
    if (a) std::memcpy(a, ha.data(), sizeof(double) * ha.size());
  }
  if (vi)
    for (long i = 0; i < 3L * c->nn; i++) vi[i] = hv[i] - 0.5 * c->dt * ha[i];
  if (m) {
    std::vector<double> hm(c->nn);
    HIPCHK(c, hipMemcpyAsync(hm.data(), c->P.m, sizeof(double) * c->nn, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < c->nn; i++) m[c->nperm[i]] = hm[i];
  }
  return c->inited ? check_flags(c) : CPI_OK;
}

extern "C" int cpi_get_increment(cpi_ctx *c, double *du) {
  if (!c || !du) return CPI_ERR_INVALID;
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return get_nodes(c, c->P.du, du);
}

extern "C" int cpi_get_element_state(cpi_ctx *c, double *sig, double *ep) {
  if (!c) return CPI_ERR_INVALID;
  if (c->phys == CPI_PHYSICS_COMPAT) {
    if (sig) std::memset(sig, 0, sizeof(double) * 6 * c->nq * (size_t)c->ne);
    if (ep) std::memset(ep, 0, sizeof(double) * c->nq * (size_t)c->ne);
    return CPI_OK;
  }
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const size_t n = (size_t)c->nq * c->ne;
  auto fetch = [&](const void *d, size_t cnt, std::vector<double> &out) -> int {
    out.resize(cnt);
    if (c->prec == CPI_PREC_F32) {
      std::vector<float> t(cnt);
      HIPCHK(c, hipMemcpyAsync(t.data(), d, 4 * cnt, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
      for (size_t i = 0; i < cnt; i++) out[i] = t[i];
    } else {
      HIPCHK(c, hipMemcpyAsync(out.data(), d, 8 * cnt, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return CPI_OK;
  };
  std::vector<double> hs, he;
  int rc;
  if (sig) {
    if ((rc = fetch(c->P.sig, 6 * n, hs))) return rc;
    for (int e = 0; e < c->ne; e++)
      for (int q = 0; q < c->nq; q++)
        for (int k = 0; k < 6; k++)
          sig[((size_t)c->eperm[e] * c->nq + q) * 6 + k] = hs[(size_t)(q * 6 + k) * c->ne + e];
  }
  if (ep) {
    if ((rc = fetch(c->P.epbar, n, he))) return rc;
    for (int e = 0; e < c->ne; e++)
      for (int q = 0; q < c->nq; q++) ep[(size_t)c->eperm[e] * c->nq + q] = he[(size_t)q * c->ne + e];
  }
  return CPI_OK;
}

extern "C" int cpi_get_stress_strain(cpi_ctx *c, double *strain, double *stress) {
  if (!c) return CPI_ERR_INVALID;
  if (!c->inited) return fail(c, CPI_ERR_STATE, "cpi_get_stress_strain before cpi_init");
  int rc = ensure_tmp(c, 12 * (size_t)c->ne);
  if (rc) return rc;
  hipError_t e = launch_stress_elem(c->P, c->d_tmp, c->stream);
  if (e != hipSuccess) return fail(c, CPI_ERR_HIP, hipGetErrorString(e));
  std::vector<double> h(12 * (size_t)c->ne), byuser(12 * (size_t)c->ne);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->d_tmp, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int e2 = 0; e2 < c->ne; e2++)
    std::memcpy(&byuser[12 * (size_t)c->eperm[e2]], &h[12 * (size_t)e2], 12 * sizeof(double));
  // element order, last element containing the node wins (dynamics_kernels.h:141-147)
  for (int e2 = 0; e2 < c->ne; e2++)
    for (int j = 0; j < c->npe; j++) {
      int n = c->conn_u[(size_t)c->npe * e2 + j];
      for (int k = 0; k < 6; k++) {
        if (strain) strain[6L * n + k] = byuser[12 * (size_t)e2 + k];
        if (stress) stress[6L * n + k] = byuser[12 * (size_t)e2 + 6 + k];
      }
    }
  return CPI_OK;
}

extern "C" int cpi_get_energy(cpi_ctx *c, double *row, int64_t *step) {
  if (!c || !row) return CPI_ERR_INVALID;
  Ctl h;
  HIPCHK(c, hipMemcpyAsync(&h, c->P.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  std::memcpy(row, h.row, sizeof(double) * CPI_NENERGY);
  if (step) *step = h.step;
  return CPI_OK;
}

extern "C" int cpi_get_energy_history(cpi_ctx *c, double *rows, int32_t cap, int32_t *count) {
  if (!c || !count || cap < 0) return CPI_ERR_INVALID;
  Ctl h;
  HIPCHK(c, hipMemcpyAsync(&h, c->P.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  const int64_t have = std::min<int64_t>(h.nrows, c->P.hist_cap);
  const int64_t n = std::min<int64_t>(have, cap);
  std::vector<double> all(8 * (size_t)c->P.hist_cap);
  HIPCHK(c, hipMemcpyAsync(all.data(), c->P.hist, sizeof(double) * all.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int64_t i = 0; i < n; i++) {
    int64_t r = h.nrows - n + i;
    std::memcpy(rows + 8 * i, &all[8 * (size_t)(r % c->P.hist_cap)], 8 * sizeof(double));
  }
  *count = (int32_t)n;
  return CPI_OK;
}

extern "C" int cpi_get_info(cpi_ctx *c, cpi_info *info) {
  if (!c || !info) return CPI_ERR_INVALID;
  std::memset(info, 0, sizeof(*info));
  info->num_nodes = c->nn;
  info->num_elements = c->ne;
  info->nodes_per_element = c->npe;
  info->num_quadrature_pts = c->nq;
  info->num_colors = c->ncol;
  info->num_blocks = c->scatter == CPI_SCATTER_BLOCK ? c->nblocks : 0;
  info->num_boundary_nodes = c->scatter == CPI_SCATTER_BLOCK ? c->plan.n_boundary : c->nn;
  info->num_partial_slots = c->scatter == CPI_SCATTER_BLOCK ? c->plan.n_slots : 0;
  // algorithmic bytes of an elastic step: connectivity 4 npe and the
  // stress read and written, 12 nq s, per element (ep is written only when
  // a point yields; the exact-return tet4 and per-point tet10 block kernels
  // read it only for points that have yielded, the other kernels read it
  // always), then per node what each pass must touch once
  const double s = c->prec == CPI_PREC_F32 ? 4 : 8;
  const bool block = c->scatter == CPI_SCATTER_BLOCK;
  const bool bits = block && c->phys == CPI_PHYSICS_EP && (c->npe == 4 || c->nq == 4);
  const double state =
      c->phys == CPI_PHYSICS_COMPAT ? 0.0 : (12.0 + (bits ? 0.0 : 1.0)) * c->nq * s;
  const double per_elem = 4.0 * c->npe + state;
  info->bytes_per_step = c->ne * per_elem + 224.0 * c->nn;
  if (block) {
    // corners as 16-bit words, or two u16 arrays
    const double nb = c->plan.n_boundary, ni = c->plan.n_interior;
    const double conn = corner_words(c) ? 2.0 * c->npe : 4.0 * c->npe;
    info->elem_kernel_bytes = c->ne * (conn + state) + ni * 104.0 + nb * 72.0;
    info->node_kernel_bytes = nb * 128.0;
  } else {
    info->elem_kernel_bytes = c->ne * per_elem + c->nn * 72.0;
    info->node_kernel_bytes = c->nn * 152.0;
  }
  info->dt = c->dt;
  info->step = c->step;
  return CPI_OK;
}

extern "C" int cpi_get_coloring(cpi_ctx *c, int32_t *color, int32_t *ncol) {
  if (!c) return CPI_ERR_INVALID;
  if (color) std::memcpy(color, c->color.data(), sizeof(int32_t) * c->ne);
  if (ncol) *ncol = c->ncol;
  return CPI_OK;
}

extern "C" int cpi_time_steps(cpi_ctx *c, int32_t nsteps, double *ms) {
  if (!c || !ms) return CPI_ERR_INVALID;
  hipEvent_t a, b;
  HIPCHK(c, hipEventCreate(&a));
  HIPCHK(c, hipEventCreate(&b));
  HIPCHK(c, hipEventRecord(a, c->stream));
  int rc = cpi_step(c, nsteps);
  HIPCHK(c, hipEventRecord(b, c->stream));
  HIPCHK(c, hipEventSynchronize(b));
  float t = 0.f;
  HIPCHK(c, hipEventElapsedTime(&t, a, b));
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  *ms = t;
  return rc;
}

extern "C" int cpi_profile_kernels(cpi_ctx *c, int32_t nsteps, cpi_kernel_time *out, int32_t cap,
                                   int32_t *count) {
  if (!c || !out || !count) return CPI_ERR_INVALID;
  if (!c->inited) return fail(c, CPI_ERR_STATE, "cpi_profile_kernels before cpi_init");
  // with the surrogate the element pass is three launches: every other step
  // times them one by one, "elem_pass" (the elastic pass), "mlp_queue" (the
  // network on the queue) and "mlp_walk" (the deferred blocks), and the
  // steps between time the three together ("elem_block": from the start of
  // the first to the end of the last).  Block-scatter kernels are timed by
  // events their launch records at the kernel's own start and end
  // (hipExtLaunchKernel); the colour and atomic scatters' launches by events
  // recorded around them.
  const bool mlp = c->scatter == CPI_SCATTER_BLOCK && c->phys == 2;
  const char *names[5] = {c->scatter == CPI_SCATTER_BLOCK ? "elem_block" : "elem_direct",
                          c->scatter == CPI_SCATTER_BLOCK ? "node_bnd" : "node_all", "elem_pass",
                          "mlp_queue", "mlp_walk"};
  double tot[5] = {0, 0, 0, 0, 0};
  int64_t launches[5] = {0, 0, 0, 0, 0};
  // ev[0], ev[1] element pass start and end; ev[2], ev[3] node pass; ev[4..9]
  // the three surrogate launches one by one (destroyed on every return, and
  // no launch keeps them)
  struct Events {
    hipEvent_t e[10] = {};
    ~Events() {
      set_launch_events(nullptr, nullptr);
      for (auto &x : e)
        if (x) (void)hipEventDestroy(x);
    }
  } evs;
  hipEvent_t *ev = evs.e;
  for (int k = 0; k < 10; k++) HIPCHK(c, hipEventCreate(&ev[k]));
  auto elapsed = [&](int a, int b, float &ms) -> int {
    HIPCHK(c, hipEventElapsedTime(&ms, ev[a], ev[b]));
    return CPI_OK;
  };
  for (int i = 0; i < nsteps; i++) {
    Params p = c->P;
    p.store_aux = 0;
    const bool split = mlp && i % 2 == 0;
    hipError_t e = hipSuccess;
    p.epart_base = 0;
    if (mlp) {
      for (int part = 1; part <= 3 && e == hipSuccess; part++) {
        if (split)
          set_launch_events(ev[2 + 2 * part], ev[3 + 2 * part]);
        else
          set_launch_events(part == 1 ? ev[0] : nullptr, part == 3 ? ev[1] : nullptr);
        e = launch_elem_block(p, c->rule, c->prec, c->phys, false, c->nblocks, c->smem, c->stream,
                              part);
      }
      set_launch_events(nullptr, nullptr);
      if (split)
        for (int k = 2; k < 5; k++) launches[k] += 1;
      else
        launches[0] += 1;
    } else if (c->scatter == CPI_SCATTER_BLOCK) {
      set_launch_events(ev[0], ev[1]);
      e = launch_elem_block(p, c->rule, c->prec, c->phys, false, c->nblocks, c->smem, c->stream);
      set_launch_events(nullptr, nullptr);
      launches[0] += 1;
    } else {
      HIPCHK(c, hipEventRecord(ev[0], c->stream));
      const bool atomic = c->scatter == CPI_SCATTER_ATOMIC;
      const int nc = atomic ? 1 : c->ncol;
      for (int k = 0; k < nc && e == hipSuccess; k++) {
        p.elem_base = atomic ? 0 : c->color_off[k];
        p.elem_count = atomic ? c->ne : c->color_off[k + 1] - c->color_off[k];
        p.epart_base = c->color_rows[k];
        e = launch_elem_direct(p, c->rule, c->prec, c->phys, atomic, c->stream);
        launches[0] += 1;
      }
      HIPCHK(c, hipEventRecord(ev[1], c->stream));
    }
    p.epart_base = c->n_elem_rows;
    if (e == hipSuccess) {
      if (c->scatter == CPI_SCATTER_BLOCK) {
        set_launch_events(ev[2], ev[3]);
        e = launch_node_bnd(p, c->prec, false, c->stream);
        set_launch_events(nullptr, nullptr);
      } else {
        HIPCHK(c, hipEventRecord(ev[2], c->stream));
        e = launch_node_all(p, false, c->stream);
        HIPCHK(c, hipEventRecord(ev[3], c->stream));
      }
    }
    launches[1] += 1;
    if (e != hipSuccess) return fail(c, CPI_ERR_HIP, hipGetErrorString(e));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float t = 0;
    int rc;
    if (split) {
      for (int k = 0; k < 3; k++) {
        if ((rc = elapsed(4 + 2 * k, 5 + 2 * k, t))) return rc;
        tot[2 + k] += t;
      }
    } else {
      if ((rc = elapsed(0, 1, t))) return rc;
      tot[0] += t;
    }
    if ((rc = elapsed(2, 3, t))) return rc;
    tot[1] += t;
  }
  c->step += nsteps;
  // one step only: the element pass is the sum of its launches
  if (launches[0] == 0 && launches[2] > 0) {
    tot[0] = tot[2] + tot[3] + tot[4];
    launches[0] = launches[2];
  }
  // the split entries only when some step was split
  int n = std::min<int>(cap, launches[2] > 0 ? 5 : 2);
  for (int k = 0; k < n; k++) {
    std::memset(out[k].name, 0, sizeof(out[k].name));
    std::strncpy(out[k].name, names[k], sizeof(out[k].name) - 1);
    out[k].total_ms = tot[k];
    out[k].launches = launches[k];
  }
  *count = n;
  return CPI_OK;
}

extern "C" int cpi_debug_strain(cpi_ctx *c, double alpha, int32_t dcase, double *W, double *V) {
  if (!c || !W || !V) return CPI_ERR_INVALID;
  if (dcase != 0 && dcase != 1) return fail(c, CPI_ERR_INVALID, "deformation case must be 0 or 1");
  const int nb = (c->ne + 255) / 256;
  int rc = ensure_tmp(c, 2 * (size_t)nb);
  if (rc) return rc;
  Params p = c->P;
  hipError_t e = launch_debug_strain(p, c->rule, alpha, dcase, c->mat.nu, c->d_tmp, c->stream);
  if (e != hipSuccess) return fail(c, CPI_ERR_HIP, hipGetErrorString(e));
  std::vector<double> h(2 * (size_t)nb);
  HIPCHK(c, hipMemcpyAsync(h.data(), c->d_tmp, sizeof(double) * h.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  double w = 0, v = 0;
  for (int i = 0; i < nb; i++) {
    w += h[2 * i];
    v += h[2 * i + 1];
  }
  *W = w;
  *V = v;
  return CPI_OK;
}

extern "C" int cpi_bandwidth_probe(int32_t device, int64_t bytes, int32_t reps, double *gbs) {
  if (!gbs || bytes < 1024 || reps < 1) {
    set_global_error("cpi_bandwidth_probe: invalid arguments");
    return CPI_ERR_INVALID;
  }
  cpi_ctx tmp;  // only for error reporting through HIPCHK
  cpi_ctx *c = &tmp;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    set_global_error("cpi_bandwidth_probe: no HIP device " + std::to_string(device));
    return CPI_ERR_NO_DEVICE;
  }
  const long n4 = (long)(bytes / 32);
  void *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = CPI_OK;
  auto run = [&]() -> int {
    HIPCHK(c, hipSetDevice(device));
    HIPCHK(c, hipMalloc(&a, n4 * 32));
    HIPCHK(c, hipMalloc(&b, n4 * 32));
    HIPCHK(c, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    HIPCHK(c, hipMemsetAsync(a, 0, n4 * 32, s));
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, launch_stream_copy(a, b, n4, s));  // warm-up
    HIPCHK(c, hipEventRecord(e0, s));
    for (int i = 0; i < reps; i++) HIPCHK(c, launch_stream_copy(a, b, n4, s));
    HIPCHK(c, hipEventRecord(e1, s));
    HIPCHK(c, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
    *gbs = 2.0 * 32.0 * (double)n4 * reps / (ms * 1e-3) / 1e9;
    return CPI_OK;
  };
  rc = run();
  if (rc) set_global_error(c->err);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

// ---------------------------------------------------------------------------
// surrogate installation, counters and return-map samples
// ---------------------------------------------------------------------------

namespace {

bool same_value(double a, double b) { return std::fabs(a - b) <= 1e-12 * std::max(std::fabs(a), std::fabs(b)); }

// re-captures the step graphs after a change of the kernel arguments
int refresh_graphs(cpi_ctx *c) {
  drop_graphs(c);
  if (c->inited && c->opt.use_graph) return capture_graphs(c);
  return CPI_OK;
}

}  // namespace

extern "C" int cpi_set_surrogate(cpi_ctx *c, const cpi_mlp *m) {
  if (!c || !m) return CPI_ERR_INVALID;
  if (c->phys != 2)
    return fail(c, CPI_ERR_STATE, "cpi_set_surrogate: the context was not created with CPI_RETURN_MLP");
  if (m->dev != c->dev) return fail(c, CPI_ERR_INVALID, "cpi_set_surrogate: network is on another device");
  if (!m->scaled) return fail(c, CPI_ERR_STATE, "cpi_set_surrogate: network has no scaling (train or set it)");
  const MlpMaterial a = mlp_material(c->mat), &b = m->mm;
  if (!same_value(a.c11, b.c11) || !same_value(a.c12, b.c12) || !same_value(a.c44, b.c44) ||
      !same_value(a.sy0, b.sy0) || !same_value(a.H, b.H))
    return fail(c, CPI_ERR_INVALID, "cpi_set_surrogate: network was made for another material");
  HIPCHK(c, hipSetDevice(c->dev));
  HIPCHK(c, hipStreamSynchronize(m->stream));
  HIPCHK(c, hipMemcpyAsync(c->mlp.d_weights, m->W, sizeof(MlpWeights), hipMemcpyDeviceToDevice, c->stream));
  MlpParams h{};
  h.weights = c->mlp.d_weights;
  h.w1 = c->mlp.d_weights->w1;
  h.w2p = c->mlp.d_weights->w2p;
  h.w3p = c->mlp.d_weights->w3p;
  h.b1 = c->mlp.d_weights->b1;
  h.b2 = c->mlp.d_weights->b2;
  h.b3 = c->mlp.d_weights->b3;
  for (int k = 0; k < kMlpIn; k++) {
    h.in_mean[k] = m->in_mean[k];
    h.in_istd[k] = 1.0f / m->in_std[k];
  }
  for (int k = 0; k < kMlpOut; k++) {
    h.out_mean[k] = m->out_mean[k];
    h.out_std[k] = m->out_std[k];
  }
  h.c11 = (float)a.c11;
  h.c12 = (float)a.c12;
  h.c44 = (float)a.c44;
  h.sy0 = (float)a.sy0;
  h.H = (float)a.H;
  HIPCHK(c, hipMemcpyAsync(c->mlp.d_params, &h, sizeof(h), hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->mlp.ready = true;
  return CPI_OK;
}

extern "C" int cpi_get_surrogate_stats(cpi_ctx *c, int64_t *accepted, int64_t *fallback) {
  if (!c) return CPI_ERR_INVALID;
  unsigned long long h[2] = {0, 0};
  if (c->mlp.d_stats) {
    std::vector<unsigned long long> all(2 * (size_t)c->mlp.nstats);
    HIPCHK(c, hipMemcpyAsync(all.data(), c->mlp.d_stats, sizeof(unsigned long long) * all.size(),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < all.size(); i += 2) {
      h[0] += all[i];
      h[1] += all[i + 1];
    }
  }
  if (accepted) *accepted = (int64_t)h[0];
  if (fallback) *fallback = (int64_t)h[1];
  return CPI_OK;
}

extern "C" int cpi_reset_surrogate_stats(cpi_ctx *c) {
  if (!c) return CPI_ERR_INVALID;
  if (c->mlp.d_stats) {
    HIPCHK(c, hipMemsetAsync(c->mlp.d_stats, 0, 2 * sizeof(unsigned long long) * c->mlp.nstats,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return CPI_OK;
}

extern "C" int cpi_samples_enable(cpi_ctx *c, int32_t cap) {
  if (!c || cap < 0) return CPI_ERR_INVALID;
  if (c->phys != CPI_PHYSICS_EP)
    return fail(c, CPI_ERR_STATE, "cpi_samples_enable: needs CPI_PHYSICS_EP with the exact return");
  HIPCHK(c, hipSetDevice(c->dev));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  int rc = mlp_samples_alloc(c->mlp, cap, c->stream, c->err);
  if (rc) return rc;
  c->P.samples = c->mlp.d_samples;
  c->P.sample_count = c->mlp.d_sample_count;
  c->P.sample_cap = c->mlp.sample_cap;
  return refresh_graphs(c);
}

extern "C" int cpi_samples_subsample(cpi_ctx *c, double keep, uint64_t seed) {
  if (!c || !(keep >= 0.0 && keep <= 1.0)) return CPI_ERR_INVALID;
  c->P.sample_keep = keep >= 1.0 ? (1ull << 32) : (unsigned long long)(keep * 4294967296.0);
  c->P.sample_seed = seed;
  return refresh_graphs(c);
}

extern "C" int cpi_samples_get(cpi_ctx *c, float *out, int32_t cap, int32_t *recorded) {
  if (!c || cap < 0 || (cap > 0 && !out)) return CPI_ERR_INVALID;
  int cnt = 0;
  if (c->mlp.d_sample_count) {
    HIPCHK(c, hipMemcpyAsync(&cnt, c->mlp.d_sample_count, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  const int k = std::min(std::min(cnt, c->mlp.sample_cap), cap);
  if (k > 0) {
    HIPCHK(c, hipMemcpyAsync(out, c->mlp.d_samples, sizeof(float) * kSampleFloats * (size_t)k,
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  if (recorded) *recorded = cnt;
  return CPI_OK;
}

extern "C" int cpi_samples_reset(cpi_ctx *c) {
  if (!c) return CPI_ERR_INVALID;
  if (c->mlp.d_sample_count) {
    HIPCHK(c, hipMemsetAsync(c->mlp.d_sample_count, 0, sizeof(int), c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  return CPI_OK;
}
