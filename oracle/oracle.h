/*
 * oracle.h -- CPU restatement of cppimpact's explicit-dynamics step.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is the checker: tests/, the smoke
 * entry point and bench.py's cpu_baseline leg load it; the product library
 * (cppimpact_amd/) never includes, links or calls it.
 *
 * Two force models are restated:
 *   compat -- the reference's step exactly as written (dynamics.h:414-512,
 *             dynamics_kernels.h:8-169): per-element K_e(x_n) from B^T D B on
 *             the 5-point rule with the reference's truncated literals,
 *             f_e = K_e * du with no stress memory, row-sum lumped mass,
 *             per-element -f_e/m accumulation, all-node wall penalty and
 *             gravity.
 *   ep     -- the elastoplastic step: updated-Lagrangian small-strain J2
 *             radial return with linear isotropic hardening, midpoint
 *             internal-energy accounting, HRZ mass for tet10, CFL dt.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   basis / quadrature / B / D / K_e : closed-form strain energies of the
 *       reference's FEA_debug known-answer cases, partition of unity,
 *       rigid-body null space (6 zero eigenvalues), patch test, K_e symmetry.
 *   row-sum mass : rho*V/4 per tet4 node; -rho*V/20 tet10 corners (closed form).
 *   HRZ mass : rho*V/36 corners, 4*rho*V/27 mid-sides (closed form).
 *   compat step : free flight v_z(t) = v0 - g t exactly, single-node wall
 *       penetration a = -d k / m, element force = K_e du against a brute-force
 *       dense K_e, leapfrog order on a 1-element mesh by hand.
 *   radial return : uniaxial-strain and pure-shear closed-form stress paths,
 *       consistency ||s|| = sqrt(2/3)(sy + H ep) after return.
 *   colouring : validity (no shared node inside a colour) + greedy minimality
 *       (every element of colour c touches colours 0..c-1), brute force on
 *       tiny meshes.
 *   CFL dt : Kuhn-tet closed form h = a/sqrt(2), and dt <= 2/omega_max of
 *       every element (dense generalized eigenvalue on the CPU).
 */
#ifndef CPPIMPACT_ORACLE_H
#define CPPIMPACT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* quadrature rules */
#define ORC_Q_REF5 0     /* tetrahedral.h:7-41, truncated literals */
#define ORC_Q_REF4 1     /* tetrahedral.h:43-74, truncated literals */
#define ORC_Q_EXACT4 2   /* degree-2 Keast rule, full precision */
#define ORC_Q_CENTROID 3 /* 1 point, w = 1/6 */

/* lumped mass kinds */
#define ORC_MASS_ROWSUM 0 /* analysis.h:136-162 */
#define ORC_MASS_HRZ 1    /* Hinton-Rock-Zienkiewicz diagonal scaling */

/* physics */
#define ORC_COMPAT 0
#define ORC_EP 1

/* energy history slots */
#define ORC_E_KINETIC 0   /* 1/2 sum m |v_n|^2 */
#define ORC_E_INTERNAL 1  /* cumulative internal work */
#define ORC_E_PLASTIC 2   /* cumulative plastic dissipation (part of 1) */
#define ORC_E_PENALTY 3   /* 1/2 k d^2 over penetrating nodes at x_n */
#define ORC_E_GRAVITY 4   /* cumulative work of gravity */
#define ORC_E_CONTACTF 5  /* total contact force along the wall normal */
#define ORC_E_BALANCE 6   /* KE + Wint + Epen - Wgrav - (KE0 + Epen0) */
#define ORC_E_MINDT 7     /* stable dt of the configuration x_{n-1} (ep) */
#define ORC_NENERGY 8

int orc_quad_npts(int rule);
double orc_quad_point(int rule, int k, double pt[3]);
void orc_basis(int npe, const double pt[3], double *N);
void orc_basis_grad(int npe, const double pt[3], double *Nxi);
double orc_inv3x3(const double A[9], double Ainv[9]);
void orc_jacobian(int npe, const double *Nxi, const double *xe, double J[9]);
void orc_bmatrix(int npe, const double Jinv[9], const double *Nxi, double *B);
void orc_dmatrix(double E, double nu, double D[36]);

/* element level (xe, due: [npe][3]) */
void orc_element_stiffness(int npe, int rule, double E, double nu,
                           const double *xe, double *Ke);
void orc_element_force_compat(int npe, int rule, double E, double nu,
                              const double *xe, const double *due,
                              double *fe);
void orc_element_mass(int npe, int rule, int kind, double rho,
                      const double *xe, double *me);
double orc_element_volume(int npe, int rule, const double *xe);
double orc_strain_energy(int npe, int rule, double E, double nu,
                         const double *xe, const double *ue);
void orc_stress_strain(int npe, double E, double nu, const double *xe,
                       const double *ue, const double pt[3], double strain[6],
                       double stress[6]);
double orc_element_dt(int npe, double E, double nu, double rho,
                      const double *xe);

/* J2 radial return on one point; returns the plastic multiplier */
double orc_radial_return(double E, double nu, double sy0, double H,
                         const double deps[6], double sig[6], double *epbar);

/* ep element force; sig/epbar are [nqp][6]/[nqp] and are updated in place.
   work[0] += internal work, work[1] += plastic dissipation. */
void orc_element_force_ep(int npe, int rule, double E, double nu, double sy0,
                          double H, const double *xe, const double *due,
                          double *sig, double *epbar, double *fe,
                          double work[2]);

/* debug_strain known-answer driver (dynamics.h:303-412) */
void orc_debug_strain(int npe, int rule, double E, double nu, int num_nodes,
                      int num_elements, const int *conn, const double *xloc,
                      double alpha, int def_case, double out[2]);

/* greedy colouring in element order, smallest free colour.
   Returns the number of colours, or -1 on allocation failure. */
int orc_greedy_coloring(int num_elements, int npe, const int *conn,
                        int num_nodes, int *color);

typedef struct orc_problem {
  int num_nodes, num_elements, npe;
  const double *X0; /* [N][3] positions after Dynamics::initialize */
  const int *conn;  /* [E][npe], 0-based */
  const double *v0; /* [N][3] */
  double E, nu, rho, sigma_y, H;
  int physics;   /* ORC_COMPAT / ORC_EP */
  int quad_rule; /* ORC_Q_* */
  int mass_kind; /* ORC_MASS_* */
  int wall_enabled, wall_dim, wall_normal;
  double wall_location, wall_stiffness;
  double gravity[3];
  double dt;
  /* per-node flags, NULL = none: contact_mask limits the wall to the flagged
     nodes (the Wall's slave set); fixed_mask holds the flagged nodes (the
     mesh's fixed set: zero initial velocity, zero acceleration, no contact).
     The reference collects both sets (mesh.h:270-329) and applies neither. */
  const unsigned char *contact_mask;
  const unsigned char *fixed_mask;
} orc_problem;

typedef struct orc_state {
  double *x;      /* [N][3] current positions */
  double *v;      /* [N][3] half-step velocity v_{n+1/2} */
  double *a;      /* [N][3] acceleration a_n */
  double *m;      /* [N]    lumped mass (one value per node) */
  double *v_i;    /* [N][3] integer-step velocity v_n */
  double *strain; /* [N][6] compat output, last element wins (may be NULL) */
  double *stress; /* [N][6] */
  double *sig;    /* [E][nqp][6] ep stress state */
  double *epbar;  /* [E][nqp] ep equivalent plastic strain */
  double *energy; /* [nsteps+1][ORC_NENERGY] (may be NULL) */
  double *Xref;   /* [N][3] scratch copy of the reference positions */
} orc_state;

/* Runs the initialisation update + half-kick, then nsteps loop iterations
   (dynamics.h:440-511).  State arrays must be allocated by the caller. */
int orc_solve(const orc_problem *p, int nsteps, orc_state *s);

/* orc_solve that also returns the wall time of the nsteps loop iterations
   (initialisation excluded); the CPU baseline uses it. */
int orc_solve_timed(const orc_problem *p, int nsteps, orc_state *s,
                    double *loop_seconds);

/* orc_solve_timed with the element and node loops split over nthreads
   POSIX threads.  Elements write their nodal contributions to per-element
   buffers and each node sums them in ascending element order, so the
   results are bit-identical to orc_solve for any thread count.  Runs the
   serial loop when nthreads <= 1 or strain/stress output is requested.
   Returns -1 on allocation failure. */
int orc_solve_threaded(const orc_problem *p, int nsteps, orc_state *s,
                       int nthreads, double *loop_seconds);

/* Stable time step of the mesh at X0 (min over elements), and the contact
   spring limit 2*sqrt(m_min/k) when k > 0 (returned in *dt_contact). */
double orc_stable_dt(const orc_problem *p, double *dt_contact);

#ifdef __cplusplus
}
#endif
#endif
