/*
 * cppimpact_hip.h -- C ABI of the MI355X (gfx950) explicit-dynamics step.
 *
 * The reference has no plugin or FFI layer: its drivers pick a backend at
 * compile time (test.cu:14-18, FEA_debug.cu:18-22) between the CPU
 * Dynamics (include/dynamics.h) and the CUDA Dynamics (include/dynamics.cuh),
 * both exposing
 *     Dynamics(Mesh*, BaseMaterial*, Wall*)          dynamics.h:64-83
 *     initialize(T pos[3], T vel[3])                 dynamics.h:99-111
 *     solve(double dt, double time_end, int export)  dynamics.h:414-512
 *                                                     dynamics.cuh:352-525
 * This library is what a third backend (CPPIMPACT_HIP_BACKEND) calls from
 * host-only C++: every entry point takes plain pointers and sizes, returns
 * 0 or a negative cpi_status, and records a message for cpi_last_error().
 * The entry point each one replaces is named beside it.
 *
 * Host arrays use the reference's layouts: positions/velocities/accelerations
 * are [num_nodes][3] doubles (Mesh::xloc, Dynamics::vel), connectivity is
 * [num_elements][nodes_per_element] 0-based int32 (Mesh::element_nodes).
 * The library keeps its own device layout (one 32-byte record per node per
 * field, stress state structure-of-arrays, elements and nodes permuted for
 * locality) and converts on upload/download.
 *
 * Thread safety: one context must not be used from two threads at once.
 */
#ifndef CPPIMPACT_HIP_H
#define CPPIMPACT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPI_ABI_VERSION 1

typedef enum cpi_status {
  CPI_OK = 0,
  CPI_ERR_INVALID = -1,     /* bad argument or unsupported combination */
  CPI_ERR_HIP = -2,         /* HIP runtime failure (gpuErrchk analogue) */
  CPI_ERR_ALLOC = -3,       /* host or device allocation failed */
  CPI_ERR_MESH = -4,        /* degenerate/inverted element, bad connectivity */
  CPI_ERR_NUMERIC = -5,     /* NaN/Inf in the state, an inverted element, or
                               a dt above the stable limit (J2 physics) */
  CPI_ERR_IO = -6,          /* file could not be opened or parsed */
  CPI_ERR_STATE = -7,       /* call made out of order (e.g. step before init) */
  CPI_ERR_NO_DEVICE = -8    /* no usable gfx950 device */
} cpi_status;

/* physics of the element force */
#define CPI_PHYSICS_COMPAT 0 /* f_e = K_e(x_n) du, no stress memory:
                                analysis.h:321-382, dynamics_kernels.h:75-148 */
#define CPI_PHYSICS_EP 1     /* f_e = sum_q w detJ B^T sigma, J2 radial return */

/* arithmetic of the element kernels (nodal state, mass, dt and energies are
   always FP64) */
#define CPI_PREC_F64 0
#define CPI_PREC_F32 1

/* nodal force assembly */
#define CPI_SCATTER_BLOCK 0  /* workgroup-local LDS assembly, deterministic */
#define CPI_SCATTER_COLOR 1  /* greedy element colouring, one pass per colour */
#define CPI_SCATTER_ATOMIC 2 /* FP64 global atomics (order not reproducible) */

/* quadrature (tetrahedral.h:7-74) */
#define CPI_QUAD_DEFAULT -1  /* compat: REF5; ep: CENTROID (tet4), EXACT4 (tet10) */
#define CPI_QUAD_REF5 0      /* the reference's 5-point rule, truncated literals */
#define CPI_QUAD_REF4 1      /* the reference's 4-point rule, truncated literals */
#define CPI_QUAD_EXACT4 2    /* 4-point degree-2 rule, full precision */
#define CPI_QUAD_CENTROID 3  /* 1 point */

/* lumped mass */
#define CPI_MASS_DEFAULT -1  /* compat: ROWSUM; ep: HRZ */
#define CPI_MASS_ROWSUM 0    /* m_i = sum_q N_i w detJ rho (analysis.h:136-162) */
#define CPI_MASS_HRZ 1       /* diagonal scaling: rho V/36, 4 rho V/27 (tet10) */

/* stress update used by CPI_PHYSICS_EP */
#define CPI_RETURN_EXACT 0   /* closed-form radial return */
#define CPI_RETURN_MLP 1     /* MLP surrogate; the exact return (FP64) where
                                its state misses a J2 condition by more than
                                mlp_tolerance * Y0 (see cpi_set_surrogate) */

/* energy row layout (cpi_get_energy) */
#define CPI_E_KINETIC 0   /* 1/2 sum m |v_n|^2 */
#define CPI_E_INTERNAL 1  /* cumulative internal work */
#define CPI_E_PLASTIC 2   /* cumulative plastic dissipation (part of INTERNAL) */
#define CPI_E_PENALTY 3   /* 1/2 k d^2 summed over penetrating nodes */
#define CPI_E_GRAVITY 4   /* cumulative work of gravity */
#define CPI_E_CONTACTF 5  /* total wall contact force along the normal */
#define CPI_E_BALANCE 6   /* KE + Wint + Epen - Wgrav - (KE0 + Epen0) */
#define CPI_E_MINDT 7     /* stable dt of the configuration stepped from */
#define CPI_NENERGY 8

typedef struct cpi_ctx cpi_ctx;

/* Mesh<T, npe> public arrays (mesh.h:14-18) */
typedef struct cpi_mesh_desc {
  int32_t num_nodes;
  int32_t num_elements;
  int32_t nodes_per_element;     /* 4 (C3D4) or 10 (C3D10 node order) */
  const double *xloc;            /* [num_nodes][3] */
  const int32_t *element_nodes;  /* [num_elements][nodes_per_element] */
  int32_t num_slave_nodes;       /* Mesh::slave_nodes; used only when
                                    cpi_options.contact_slave_only is set */
  const int32_t *slave_nodes;
  int32_t num_fixed_nodes;       /* Mesh::fixed_nodes (mesh.h:270-298); used
                                    only when cpi_options.fixed_bc is set */
  const int32_t *fixed_nodes;
} cpi_mesh_desc;

/* BaseMaterial{E, rho, nu} + Elastoplastic{beta, H, Y0}
   (basematerial.h:5-20, elastoplastic.h:9-22).  beta is accepted and not
   used (the reference never reads it); hardening is linear isotropic. */
typedef struct cpi_material {
  double E, rho, nu, beta, H, Y0;
} cpi_material;

/* Wall<T, dim, Basis>(name, location, stiffness, slave, n_slave, normal)
   (wall.h:13-44): axis-aligned rigid plane; penetration d < 0 adds
   -d * normal * stiffness / m to the acceleration along dim. */
typedef struct cpi_wall {
  int32_t enabled;
  int32_t dim;      /* 0, 1, 2 (the reference instantiates 2) */
  int32_t normal;   /* +1 or -1 */
  double location;
  double stiffness;
} cpi_wall;

typedef struct cpi_options {
  int32_t physics;          /* CPI_PHYSICS_* */
  int32_t precision;        /* CPI_PREC_* */
  int32_t scatter;          /* CPI_SCATTER_* */
  int32_t quadrature;       /* CPI_QUAD_* */
  int32_t mass;             /* CPI_MASS_* */
  int32_t stress_update;    /* CPI_RETURN_* */
  int32_t use_graph;        /* replay the step as a hipGraph (default 1) */
  int32_t track_energy;     /* per-step energy rows (default 1) */
  int32_t contact_slave_only; /* 0 = every node, as the reference does */
  int32_t block_elements;   /* elements per workgroup for SCATTER_BLOCK, 0 = auto
                               (384 tet4, 128 tet10); at most 384 for tet4 and
                               128 for tet10 on a 4-point rule */
  int32_t device;           /* HIP device ordinal */
  int32_t fixed_bc;         /* 1: hold the mesh's fixed nodes (zero initial
                               velocity, zero acceleration, no contact);
                               0 = none, as the reference (which collects
                               the set and never applies it) */
  int32_t debug_fill_nan;   /* 1: fill every device buffer of the context with
                               all-ones bytes (NaN as a float or double, -1 as
                               an integer) when it is allocated, so a read of
                               memory the context never wrote shows as a
                               non-finite state (diagnostics; default 0) */
  int32_t stream_state;     /* the block kernels' stress-state accesses
                               (tet4 and the 4-point tet10 kernel):
                               0 = non-temporal when a step's working set
                               (state, node records, partial forces,
                               connectivity) exceeds 192 MB (default),
                               1 = always, 2 = never; results are the
                               same bits.  tet4 plans with a node in more
                               than 128 elements of one block keep the
                               state cached */
  double gravity[3];        /* default (0, 0, -9.81): dynamics_kernels.h:165-167 */
  double mlp_tolerance;     /* default 1e-3 */
} cpi_options;

typedef struct cpi_info {
  int32_t num_nodes, num_elements, nodes_per_element, num_quadrature_pts;
  int32_t num_colors;       /* greedy colouring (always computed) */
  int32_t num_blocks;       /* SCATTER_BLOCK workgroups */
  int32_t num_boundary_nodes; /* nodes shared between workgroups */
  int32_t num_partial_slots;
  double bytes_per_step;    /* algorithmic HBM bytes of one step */
  double elem_kernel_bytes; /* algorithmic bytes of the element kernel */
  double node_kernel_bytes; /* algorithmic bytes of the node kernel */
  double dt;                /* current time step */
  int64_t step;             /* completed loop iterations since cpi_init */
} cpi_info;

typedef struct cpi_kernel_time {
  char name[48];
  double total_ms;
  int64_t launches;
} cpi_kernel_time;

/* ---- context ----------------------------------------------------------- */

void cpi_default_options(cpi_options *opt);

/* Dynamics ctor + the device half of Dynamics<...>::allocate()
   (dynamics.h:64-83, dynamics.cuh:285-335).  Copies the mesh, builds the
   colouring and the workgroup partition, uploads everything. */
int cpi_create(const cpi_mesh_desc *mesh, const cpi_material *mat,
               const cpi_wall *wall, const cpi_options *opt, cpi_ctx **out);

/* Dynamics::~Dynamics + deallocate() (dynamics.h:85-95, dynamics.cuh:337-350) */
void cpi_destroy(cpi_ctx *ctx);

/* Message of the last failure on ctx (or of the last failed cpi_create when
   ctx is NULL).  Never NULL. */
const char *cpi_last_error(const cpi_ctx *ctx);

/* Dynamics::initialize (dynamics.h:99-111): reference positions X0 and
   initial velocities, [num_nodes][3].  Either may be NULL (keep mesh xloc /
   zero velocity). */
int cpi_set_state(cpi_ctx *ctx, const double *x0, const double *v0);

/* Stable time step of the configuration X0: element CFL limit and the wall
   penalty spring limit 2 sqrt(m_min / k).  Needs the mass, so it assembles
   it if cpi_init has not run. */
int cpi_stable_dt(cpi_ctx *ctx, double *dt_element, double *dt_contact);

/* cpi_stable_dt of the current configuration for callers that cannot
   allocate (serving loops and harnesses that forbid allocation and check
   the stable step between cpi_step calls; it waits for the context's
   stream, so it is not a call to capture).  No counterpart in the
   reference.
   It calls no hipMalloc / hipFree and no allocating routine: its partial
   minima go to a file-scope device array of 65,538 doubles, one per GPU,
   held until the process ends.  Needs cpi_init (the mass); meshes above
   16,777,216 elements are refused with CPI_ERR_INVALID.  Calls that overlap
   on one GPU share that array and race.  Same results as cpi_stable_dt. */
int cpi_stable_dt_noalloc(cpi_ctx *ctx, double *dt_element, double *dt_contact);

/* First half of Dynamics::solve (dynamics.h:439-466): x = X0, mass assembly,
   the initial update with du = 0, and the half-kick v += dt/2 a0.
   dt <= 0 selects dt = -dt * dt_stable (e.g. -0.5 gives half the stable
   step; dt == 0 gives 0.5 * dt_stable). */
int cpi_init(cpi_ctx *ctx, double dt);

/* nsteps iterations of the loop body of Dynamics::solve
   (dynamics.h:471-511, dynamics.cuh:450-517): du = dt v; update(); x += du;
   v += dt a; v_i = v - dt/2 a.  Asynchronous: returns after enqueueing. */
int cpi_step(cpi_ctx *ctx, int32_t nsteps);

/* Waits for queued work; reports CPI_ERR_NUMERIC if a NaN/Inf or an inverted
   element was seen, or, with CPI_PHYSICS_EP, if the fixed dt exceeded the
   stable limit of a configuration stepped from (the element CFL limit,
   reduced on the device on the last step of every cpi_step call, or the
   wall spring's 2 sqrt(m_min / k)).  The reference steps with a fixed dt
   and no such check (test.cu:83); dt is never changed here either. */
int cpi_synchronize(cpi_ctx *ctx);

/* Host copies of Dynamics::global_xloc, vel (v_{n+1/2}), global_acc (a_n),
   global_mass (one value per node) and vel_i (v_n).  Any pointer may be NULL. */
int cpi_get_state(cpi_ctx *ctx, double *x, double *v, double *a, double *m,
                  double *v_i);

/* du of the last completed step [num_nodes][3] (Dynamics::global_dof,
   dynamics.h:473-476); zeros before the first step. */
int cpi_get_increment(cpi_ctx *ctx, double *du);

/* Quadrature-point state: sig [num_elements][nqp][6] (xx yy zz xy yz xz),
   epbar [num_elements][nqp].  Either may be NULL. */
int cpi_get_element_state(cpi_ctx *ctx, double *sig, double *epbar);

/* Nodal strain/stress as the reference's update() leaves them
   (dynamics_kernels.h:117-147): per element at xi = 0 from the total
   displacement x_n + du - X0 evaluated with B(x_n), each node taking the
   value of the highest-numbered element that contains it.  Computed for the
   last completed step.  [num_nodes][6] each. */
int cpi_get_stress_strain(cpi_ctx *ctx, double *strain, double *stress);

/* Latest energy row (CPI_E_*) and the step it belongs to. */
int cpi_get_energy(cpi_ctx *ctx, double *row, int64_t *step);

/* Up to cap most recent rows, oldest first; *count receives the number. */
int cpi_get_energy_history(cpi_ctx *ctx, double *rows, int32_t cap,
                           int32_t *count);

int cpi_get_info(cpi_ctx *ctx, cpi_info *info);

/* Greedy colour of every element (user order) and the colour count. */
int cpi_get_coloring(cpi_ctx *ctx, int32_t *color, int32_t *num_colors);

/* Enqueues nsteps steps between two HIP events on the context's stream and
   returns the elapsed GPU time. */
int cpi_time_steps(cpi_ctx *ctx, int32_t nsteps, double *elapsed_ms);

/* Runs nsteps steps launching each kernel separately on the context's
   stream and times it with HIP events: block-scatter kernels by events
   their launch records at the kernel's own start and end
   (hipExtLaunchKernel), the colour and atomic scatters by events recorded
   around their launches.  Fills per-kernel totals: the element pass
   ("elem_block" / "elem_direct"; in the surrogate mode from the start of
   its first launch to the end of its third) and the node pass ("node_bnd" /
   "node_all"), then, in the surrogate mode, the element pass's three
   launches one by one on every other step ("elem_pass", "mlp_queue",
   "mlp_walk").  *count = entries written (at most cap). */
int cpi_profile_kernels(cpi_ctx *ctx, int32_t nsteps, cpi_kernel_time *out,
                        int32_t cap, int32_t *count);

/* Dynamics::debug_strain on the GPU (dynamics.h:303-412): applies the
   known-answer displacement field (case 0: alpha x (1,-nu,-nu); case 1:
   alpha x^2/2 (1,-nu,-nu)) and returns total strain energy and volume with
   the context's quadrature. */
int cpi_debug_strain(cpi_ctx *ctx, double alpha, int32_t def_case,
                     double *energy, double *volume);

/* STREAM-style copy of bytes (read + write counted) repeated reps times on
   device; returns the achieved GB/s.  No counterpart in the reference: it
   measures the HBM bandwidth the roofline is compared against and
   calibrates the memory counters (a copy moves a known number of bytes). */
int cpi_bandwidth_probe(int32_t device, int64_t bytes, int32_t reps, double *gbs);

/* ---- MLP surrogate of the J2 return map --------------------------------- *
 * No counterpart in the reference (its stress update is the linear
 * K_e du of dynamics_kernels.h:101-114).  A 13 -> 64 -> 64 -> 7 SiLU network
 * evaluated with BF16 MFMA (FP32 accumulation) replaces the radial return
 * at yielding points of CPI_PHYSICS_EP + CPI_RETURN_MLP contexts (tet4,
 * FP32, SCATTER_BLOCK).  Its state is accepted only where it keeps the
 * pressure, does not turn the trial deviator, lowers the deviator's norm by
 * sqrt(6) G dep (consistency) and lies on the yield surface, each within
 * mlp_tolerance * Y0; elsewhere the point takes the exact radial return in
 * FP64.  A step queues the yielding points of the blocks whose trial stress
 * yields, evaluates the network on the whole queue in one launch and
 * finishes those blocks with its outputs.
 *
 * Inputs (standardized): strain increment [6], stress_n [6], eq. plastic
 * strain_n.  Outputs (standardized): (sigma_trial - sigma_{n+1}) / q [6] and
 * Y0 * dep / q, with q = |dev sigma_trial| - sqrt(2/3) (Y0 + H ep_n) the
 * trial overshoot.  Parameters, canonical order (5511 floats):
 * W1[64][13], b1[64], W2[64][64], b2[64], W3[7][64], b3[7].
 *
 * A sample, as recorded by cpi_samples_enable, is 20 floats:
 * de[6], sigma_n[6], ep_n, sigma_{n+1}[6], ep_{n+1}. */
#define CPI_MLP_INPUTS 13
#define CPI_MLP_OUTPUTS 7
#define CPI_MLP_NPARAMS 5511
#define CPI_SAMPLE_FLOATS 20

/* outcome of one point of cpi_mlp_return_map */
#define CPI_MLP_ELASTIC 0
#define CPI_MLP_ACCEPTED 1
#define CPI_MLP_FALLBACK 2

typedef struct cpi_mlp cpi_mlp;

typedef struct cpi_mlp_train_opts {
  int32_t epochs;       /* default 60 */
  int32_t batch;        /* samples per Adam step, default 1024, <= 65536 */
  float lr, lr_final;   /* cosine decay from lr to lr_final */
  float beta1, beta2, eps;
  uint64_t seed;        /* shuffling */
} cpi_mlp_train_opts;

void cpi_mlp_default_train_opts(cpi_mlp_train_opts *opt);

/* Network tied to a material (D and the yield curve), Glorot-uniform
   weights from SplitMix64(seed), zero biases, identity scaling. */
int cpi_mlp_create(int32_t device, uint64_t seed, const cpi_material *mat,
                   cpi_mlp **out);
void cpi_mlp_destroy(cpi_mlp *m);
const char *cpi_mlp_last_error(const cpi_mlp *m);

/* FP32 master parameters, canonical order; setting them clears Adam. */
int cpi_mlp_set_params(cpi_mlp *m, const float *params);
int cpi_mlp_get_params(cpi_mlp *m, float *params);

/* standardization: x_std = (x - mean) / std, per feature and output */
int cpi_mlp_set_scaling(cpi_mlp *m, const float *in_mean, const float *in_std,
                        const float *out_mean, const float *out_std);
int cpi_mlp_get_scaling(cpi_mlp *m, float *in_mean, float *in_std,
                        float *out_mean, float *out_std);
/* mean / std of the samples with q > 1e-3 Y0 */
int cpi_mlp_fit_scaling(cpi_mlp *m, const float *samples, int32_t n);
/* standardized (x[kept][13], t[kept][7]) of the samples with q > 1e-3 Y0 */
int cpi_mlp_targets(cpi_mlp *m, const float *samples, int32_t n, float *x,
                    float *t, int32_t *kept);

/* y[n][7] = network(x[n][13]), standardized, through the inference kernel
   the element pass uses. */
int cpi_mlp_forward(cpi_mlp *m, const float *x, int32_t n, float *y);

/* Gradient of L = 1/(2n) sum |y - t|^2 over the parameters (canonical
   order) by the BF16-MFMA training kernel; n <= 65536. */
int cpi_mlp_gradient(cpi_mlp *m, const float *x, const float *t, int32_t n,
                     float *grad, float *loss);

/* Adam on samples (scaling fitted first if never set); loss_history gets
   the mean loss of each epoch (may be NULL). */
int cpi_mlp_train(cpi_mlp *m, const float *samples, int32_t n,
                  const cpi_mlp_train_opts *opt, float *loss_history);

/* The element pass's surrogate return map on n points (FP32): trial
   stress sigma_n + D de, then the surrogate's state where it passes the
   acceptance conditions above, else the exact FP64 return; how[i] is
   CPI_MLP_ELASTIC / ACCEPTED / FALLBACK. */
int cpi_mlp_return_map(cpi_mlp *m, double tolerance, int32_t n,
                       const float *de, const float *sigma_n,
                       const float *ep_n, float *sigma_out, float *ep_out,
                       int32_t *how);

/* Installs a copy of m's weights and scaling into a CPI_RETURN_MLP context
   (the material must match); required before cpi_init. */
int cpi_set_surrogate(cpi_ctx *ctx, const cpi_mlp *m);
/* points that took the surrogate / fell back to the exact return since the
   last reset */
int cpi_get_surrogate_stats(cpi_ctx *ctx, int64_t *accepted, int64_t *fallback);
int cpi_reset_surrogate_stats(cpi_ctx *ctx);

/* Records (de, sigma_n, ep_n, sigma_{n+1}, ep_{n+1}) of every yielding
   point of a CPI_PHYSICS_EP + CPI_RETURN_EXACT context into a device buffer
   of capacity samples (0 disables); points past the capacity are counted
   and dropped. */
int cpi_samples_enable(cpi_ctx *ctx, int32_t capacity);
/* Keeps only a fixed subset of the yielding points: point q of element e in
   step n is recorded when a 32-bit hash of (seed, n, e, q) falls below
   keep * 2^32 (keep = 1, the default, records every point).  The subset
   depends on the run alone, not on the order of the atomic slot counter,
   so a training set drawn from long runs is reproducible; the counter then
   counts the kept points. */
int cpi_samples_subsample(cpi_ctx *ctx, double keep, uint64_t seed);
/* copies min(recorded, capacity, cap) samples; *recorded = points seen */
int cpi_samples_get(cpi_ctx *ctx, float *out, int32_t cap, int32_t *recorded);
int cpi_samples_reset(cpi_ctx *ctx);

/* ---- host-only utilities (no GPU needed) -------------------------------- */

/* Greedy colouring in element order, smallest free colour; returns the
   number of colours (>= 1) or a negative status. */
int cpi_greedy_coloring(int32_t num_elements, int32_t nodes_per_element,
                        const int32_t *element_nodes, int32_t num_nodes,
                        int32_t *color);

/* Structured box of nx*ny*nz hexes, each split into 6 tets sharing the cell
   diagonal (conforming).  Node ids run x fastest; tet10 mid-side nodes sit
   at edge midpoints on the doubled lattice.  Interior nodes are jittered by
   U(-jitter, jitter) * spacing per coordinate from SplitMix64(seed).
   Call with xloc == NULL to get sizes only. */
typedef struct cpi_box_spec {
  int32_t nx, ny, nz;
  int32_t quadratic;
  double lx, ly, lz;
  double origin[3];
  double jitter;
  uint64_t seed;
} cpi_box_spec;
int cpi_mesh_box(const cpi_box_spec *spec, int32_t *num_nodes,
                 int32_t *num_elements, double *xloc, int32_t *element_nodes);

/* Thin-walled cylinder about z: (theta, z, r) grid of nt x nz x nr cells,
   periodic in theta, mid-surface radius R, length L, thickness t; cells split
   like the box; tet10 mid-side nodes at straight-edge midpoints. */
typedef struct cpi_cylinder_spec {
  int32_t ntheta, nz, nr;
  int32_t quadratic;
  double radius, length, thickness;
  double origin[3];
} cpi_cylinder_spec;
int cpi_mesh_cylinder(const cpi_cylinder_spec *spec, int32_t *num_nodes,
                      int32_t *num_elements, double *xloc,
                      int32_t *element_nodes);

/* Mesh::load_mesh (mesh.h:57-267) for the .inp subset the reference reads:
   *Node, *Element (nodes_per_element ids taken per line), *Nset.  Ids are
   1-based and assumed contiguous, as in the reference.  Node sets are
   returned in name order; slave/fixed sets are those whose name contains
   "slave"/"fixed" (case-insensitive).  Two-phase: call with out arrays NULL
   to get the counts.  Returns CPI_ERR_IO when the file cannot be opened. */
typedef struct cpi_inp_counts {
  int32_t num_nodes, num_elements, num_node_sets, num_set_indices;
  int32_t num_fixed_nodes, num_slave_nodes, names_bytes;
} cpi_inp_counts;
int cpi_inp_read(const char *path, int32_t nodes_per_element,
                 cpi_inp_counts *counts, double *xloc, int32_t *element_nodes,
                 int32_t *node_set_starts, int32_t *node_set_indices,
                 int32_t *fixed_nodes, int32_t *slave_nodes, char *names);

/* Writes an .inp the reader above (and the reference's) accepts: *Node,
   *Element (C3D4/C3D10), and Nset=fixed / Nset=slave holding the given
   nodes (each only when not empty). */
int cpi_inp_write(const char *path, const cpi_mesh_desc *mesh);

#ifdef __cplusplus
}
#endif
#endif /* CPPIMPACT_HIP_H */
