/*
 * oracle.c -- CPU restatement of cppimpact's explicit-dynamics step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, the smoke entry point and
 * bench.py's cpu_baseline leg.  The product (cppimpact_amd/) never uses it.
 *
 * Written from the reference's algorithm, in the reference's operation order
 * where that decides rounding (the reference is built as ISO C++17, so GCC
 * does not contract a*b+c into FMAs; this file is compiled with
 * -ffp-contract=off for the same reason).  Each function names the reference
 * lines it restates.  See oracle.h for what pins each function.
 */
#define _POSIX_C_SOURCE 199309L
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ---------------------------------------------------------------------- */
/* quadrature                                                              */
/* ---------------------------------------------------------------------- */

int orc_quad_npts(int rule) {
  switch (rule) {
    case ORC_Q_REF5: return 5;
    case ORC_Q_REF4: return 4;
    case ORC_Q_EXACT4: return 4;
    case ORC_Q_CENTROID: return 1;
    default: return 0;
  }
}

/* ORC_Q_REF5 restates tetrahedral.h:12-40 (weights -0.13333333333 and 0.075,
   off-centre coordinate 0.16666666666 exactly as the reference writes them);
   ORC_Q_REF4 restates tetrahedral.h:48-73. */
double orc_quad_point(int rule, int k, double pt[3]) {
  if (rule == ORC_Q_REF5) {
    const double s = 0.16666666666;
    switch (k) {
      case 0: pt[0] = 0.25; pt[1] = 0.25; pt[2] = 0.25; return -0.13333333333;
      case 1: pt[0] = s; pt[1] = s; pt[2] = s; return 0.075;
      case 2: pt[0] = 0.5; pt[1] = s; pt[2] = s; return 0.075;
      case 3: pt[0] = s; pt[1] = 0.5; pt[2] = s; return 0.075;
      case 4: pt[0] = s; pt[1] = s; pt[2] = 0.5; return 0.075;
    }
  } else if (rule == ORC_Q_REF4 || rule == ORC_Q_EXACT4) {
    double al, be, w;
    if (rule == ORC_Q_REF4) {
      al = 0.58541020; be = 0.13819660; w = 0.041666667;
    } else {
      al = (5.0 + 3.0 * sqrt(5.0)) / 20.0;
      be = (5.0 - sqrt(5.0)) / 20.0;
      w = 1.0 / 24.0;
    }
    pt[0] = be; pt[1] = be; pt[2] = be;
    if (k >= 1 && k <= 3) pt[k - 1] = al;
    return (k >= 0 && k <= 3) ? w : 0.0;
  } else if (rule == ORC_Q_CENTROID) {
    pt[0] = pt[1] = pt[2] = 0.25;
    return k == 0 ? 1.0 / 6.0 : 0.0;
  }
  pt[0] = pt[1] = pt[2] = 0.0;
  return 0.0;
}

/* ---------------------------------------------------------------------- */
/* basis                                                                   */
/* ---------------------------------------------------------------------- */

/* Linear: tetrahedral.h:521-549 (N0 = 1-xi-eta-zeta, N1 = xi, N2 = eta,
   N3 = zeta).  Quadratic: tetrahedral.h:182-250 -- corners 0,1,2,3 sit at
   barycentric L2=xi, L3=eta, L1=1-xi-eta-zeta, L4=zeta; mid-sides are
   4:(0,1) 5:(1,2) 6:(0,2) 7:(0,3) 8:(1,3) 9:(2,3) (Abaqus C3D10 order). */
void orc_basis(int npe, const double p[3], double *N) {
  if (npe == 4) {
    N[0] = 1.0 - p[0] - p[1] - p[2];
    N[1] = p[0];
    N[2] = p[1];
    N[3] = p[2];
    return;
  }
  double L1 = 1.0 - p[0] - p[1] - p[2], L2 = p[0], L3 = p[1], L4 = p[2];
  N[0] = (2 * L2 - 1) * L2;
  N[1] = (2 * L3 - 1) * L3;
  N[2] = (2 * L1 - 1) * L1;
  N[3] = (2 * L4 - 1) * L4;
  N[4] = 4 * L2 * L3;
  N[5] = 4 * L1 * L3;
  N[6] = 4 * L1 * L2;
  N[7] = 4 * L2 * L4;
  N[8] = 4 * L3 * L4;
  N[9] = 4 * L1 * L4;
}

/* Nxi[3*i + q] = dN_i / dxi_q */
void orc_basis_grad(int npe, const double p[3], double *G) {
  if (npe == 4) {
    static const double g[12] = {-1, -1, -1, 1, 0, 0, 0, 1, 0, 0, 0, 1};
    memcpy(G, g, sizeof(g));
    return;
  }
  double x = p[0], y = p[1], z = p[2];
  G[0] = 4 * x - 1; G[1] = 0.0; G[2] = 0.0;
  G[3] = 0.0; G[4] = 4 * y - 1; G[5] = 0.0;
  /* the reference sums the three terms in three different orders */
  G[6] = 4 * x + 4 * y + 4 * z - 3;
  G[7] = 4 * y + 4 * x + 4 * z - 3;
  G[8] = 4 * z + 4 * x + 4 * y - 3;
  G[9] = 0.0; G[10] = 0.0; G[11] = 4.0 * z - 1.0;
  G[12] = 4 * y; G[13] = 4 * x; G[14] = 0.0;
  G[15] = -4 * y; G[16] = -8 * y - 4 * x - 4 * z + 4; G[17] = -4 * y;
  G[18] = -8 * x - 4 * y - 4 * z + 4; G[19] = -4 * x; G[20] = -4 * x;
  G[21] = 4 * z; G[22] = 0.0; G[23] = 4 * x;
  G[24] = 0.0; G[25] = 4 * z; G[26] = 4 * y;
  G[27] = -4 * z; G[28] = -4 * z; G[29] = -8 * z - 4 * x - 4 * y + 4;
}

/* J[p][q] = sum_i dN_i/dxi_q * x_{i,p}  (tetrahedral.h:252-269, 551-567) */
void orc_jacobian(int npe, const double *G, const double *xe, double J[9]) {
  for (int a = 0; a < 9; a++) J[a] = 0.0;
  for (int p = 0; p < 3; p++)
    for (int q = 0; q < 3; q++)
      for (int i = 0; i < npe; i++) J[p * 3 + q] += G[3 * i + q] * xe[3 * i + p];
}

/* cofactor inverse, det expanded along the last row, 1/det multiplied in
   (physics.h:38-58) */
double orc_inv3x3(const double A[9], double Ai[9]) {
  double det = (A[8] * (A[0] * A[4] - A[3] * A[1]) -
                A[7] * (A[0] * A[5] - A[3] * A[2]) +
                A[6] * (A[1] * A[5] - A[2] * A[4]));
  double r = 1.0 / det;
  Ai[0] = (A[4] * A[8] - A[5] * A[7]) * r;
  Ai[1] = -(A[1] * A[8] - A[2] * A[7]) * r;
  Ai[2] = (A[1] * A[5] - A[2] * A[4]) * r;
  Ai[3] = -(A[3] * A[8] - A[5] * A[6]) * r;
  Ai[4] = (A[0] * A[8] - A[2] * A[6]) * r;
  Ai[5] = -(A[0] * A[5] - A[2] * A[3]) * r;
  Ai[6] = (A[3] * A[7] - A[4] * A[6]) * r;
  Ai[7] = -(A[0] * A[7] - A[1] * A[6]) * r;
  Ai[8] = (A[0] * A[4] - A[1] * A[3]) * r;
  return det;
}

static void dndx_of(int npe, const double Ji[9], const double *G, double *dN) {
  for (int n = 0; n < npe; n++)
    for (int i = 0; i < 3; i++) {
      double s = 0.0;
      for (int j = 0; j < 3; j++) s += Ji[j * 3 + i] * G[n * 3 + j];
      dN[3 * n + i] = s;
    }
}

/* B is [6][3*npe] row-major, caller zeroes it.  Voigt rows xx, yy, zz,
   gamma_xy, gamma_yz, gamma_xz (tetrahedral.h:442-484, 597-637). */
void orc_bmatrix(int npe, const double Ji[9], const double *G, double *B) {
  const int n3 = 3 * npe;
  double dN[30];
  dndx_of(npe, Ji, G, dN);
  for (int n = 0; n < npe; n++) {
    const double *d = dN + 3 * n;
    int c = 3 * n;
    B[0 * n3 + c + 0] = d[0];
    B[1 * n3 + c + 1] = d[1];
    B[2 * n3 + c + 2] = d[2];
    B[3 * n3 + c + 0] = d[1];
    B[3 * n3 + c + 1] = d[0];
    B[4 * n3 + c + 1] = d[2];
    B[4 * n3 + c + 2] = d[1];
    B[5 * n3 + c + 0] = d[2];
    B[5 * n3 + c + 2] = d[0];
  }
}

/* isotropic elasticity, entries scaled by E/((1+nu)(1-2nu)) after filling
   (tetrahedral.h:486-512) */
void orc_dmatrix(double E, double nu, double D[36]) {
  for (int i = 0; i < 36; i++) D[i] = 0.0;
  D[0] = D[7] = D[14] = 1 - nu;
  D[21] = D[28] = D[35] = (1 - 2 * nu) / 2;
  D[1] = D[6] = nu;
  D[2] = D[12] = nu;
  D[8] = D[13] = nu;
  for (int i = 0; i < 36; i++) D[i] *= E / ((1 + nu) * (1 - 2 * nu));
}

/* dense B^T D B in the reference's (k, l, i, j) loop order
   (analysis.h:293-319) */
static inline void btdb(int n, const double *B, const double *D, double *K) {
  for (int k = 0; k < 6; ++k) {
    const double *Bk = B + k * n;
    for (int l = 0; l < 6; ++l) {
      double Dkl = D[k * 6 + l];
      const double *Bl = B + l * n;
      for (int i = 0; i < n; ++i) {
        double t = Bk[i] * Dkl;
        double *row = K + i * n;
        for (int j = 0; j < n; ++j) row[j] += t * Bl[j];
      }
    }
  }
}

/* y accumulates: y[i] += beta; y[i] += alpha*A[i][j]*x[j]
   (cppimpact_blas.h:9-28, called with beta = 0) */
static inline void gemv_acc(int m, int n, const double *A, const double *x, double *y) {
  for (int i = 0; i < m; i++) {
    y[i] += 0.0;
    for (int j = 0; j < n; j++) y[i] += 1.0 * A[i * n + j] * x[j];
  }
}

/* ---------------------------------------------------------------------- */
/* element level                                                           */
/* ---------------------------------------------------------------------- */

/* K_e = sum_q (w_q detJ_q) B^T D B  (analysis.h:321-378).  Instantiated
   with the matrix size as a constant (12 or 30), as the reference's
   templates are, so the compiler can unroll and vectorise the loops. */
#define ORC_STIFFNESS(NAME, NPE)                                          \
  static void NAME(int rule, double E, double nu, const double *xe,       \
                   double *K) {                                           \
    enum { n = 3 * NPE };                                                 \
    double pt[3], G[30], J[9], Ji[9], B[6 * n], D[36], BDB[n * n];        \
    memset(K, 0, sizeof(double) * n * n);                                 \
    orc_dmatrix(E, nu, D);                                                \
    for (int q = 0; q < orc_quad_npts(rule); q++) {                       \
      double w = orc_quad_point(rule, q, pt);                             \
      orc_basis_grad(NPE, pt, G);                                         \
      orc_jacobian(NPE, G, xe, J);                                        \
      double detJ = orc_inv3x3(J, Ji);                                    \
      memset(B, 0, sizeof(B));                                            \
      orc_bmatrix(NPE, Ji, G, B);                                         \
      memset(BDB, 0, sizeof(BDB));                                        \
      btdb(n, B, D, BDB);                                                 \
      for (int j = 0; j < n * n; j++) K[j] += w * detJ * BDB[j];          \
    }                                                                     \
  }
ORC_STIFFNESS(stiffness4, 4)
ORC_STIFFNESS(stiffness10, 10)

void orc_element_stiffness(int npe, int rule, double E, double nu,
                           const double *xe, double *K) {
  if (npe == 4)
    stiffness4(rule, E, nu, xe, K);
  else
    stiffness10(rule, E, nu, xe, K);
}

/* f_e = K_e(x_n) * du  (analysis.h:321-382) */
void orc_element_force_compat(int npe, int rule, double E, double nu,
                              const double *xe, const double *due,
                              double *fe) {
  const int n = 3 * npe;
  double K[900];
  orc_element_stiffness(npe, rule, E, nu, xe, K);
  memset(fe, 0, sizeof(double) * n);
  gemv_acc(n, n, K, due, fe);
}

double orc_element_volume(int npe, int rule, const double *xe) {
  double pt[3], G[30], J[9], Ji[9], V = 0.0;
  for (int q = 0; q < orc_quad_npts(rule); q++) {
    double w = orc_quad_point(rule, q, pt);
    orc_basis_grad(npe, pt, G);
    orc_jacobian(npe, G, xe, J);
    V += w * orc_inv3x3(J, Ji);
  }
  return V;
}

/* ROWSUM: m_i = sum_q N_i w_q detJ_q rho (analysis.h:136-162).
   HRZ: consistent-mass diagonal scaled to the element mass; for straight
   tets this is rho V/4 (tet4) and rho V/36 corner, 4 rho V/27 mid-side
   (tet10).  The diagonal is integrated with the exact 4-point rule, which is
   exact for the degree-4 products N_i^2 of tet10 only up to degree 2, so the
   closed form is used for the diagonal ratios instead. */
void orc_element_mass(int npe, int rule, int kind, double rho,
                      const double *xe, double *me) {
  if (kind == ORC_MASS_ROWSUM) {
    double pt[3], G[30], J[9], Ji[9], N[10];
    for (int i = 0; i < npe; i++) {
      double mi = 0.0;
      for (int q = 0; q < orc_quad_npts(rule); q++) {
        double w = orc_quad_point(rule, q, pt);
        orc_basis_grad(npe, pt, G);
        orc_jacobian(npe, G, xe, J);
        double detJ = orc_inv3x3(J, Ji);
        orc_basis(npe, pt, N);
        mi += N[i] * w * detJ * rho;
      }
      me[i] = mi;
    }
    return;
  }
  double M = rho * orc_element_volume(npe, ORC_Q_CENTROID, xe);
  if (npe == 4) {
    for (int i = 0; i < 4; i++) me[i] = 0.25 * M;
  } else {
    for (int i = 0; i < 4; i++) me[i] = M / 36.0;
    for (int i = 4; i < 10; i++) me[i] = 4.0 * M / 27.0;
  }
}

/* W_e = 1/2 u^T K_e u (analysis.h:384-453).  The reference inverts the
   partition-of-unity Jacobian here; for straight-edged elements with
   mid-side nodes at edge midpoints that map equals the standard one. */
double orc_strain_energy(int npe, int rule, double E, double nu,
                         const double *xe, const double *ue) {
  const int n = 3 * npe;
  double K[900], Ku[30];
  orc_element_stiffness(npe, rule, E, nu, xe, K);
  memset(Ku, 0, sizeof(Ku));
  gemv_acc(n, n, K, ue, Ku);
  double W = 0.0;
  for (int j = 0; j < n; j++) W += 0.5 * ue[j] * Ku[j];
  return W;
}

/* strain = B(pt) u, stress = D strain (analysis.h:479-503) */
void orc_stress_strain(int npe, double E, double nu, const double *xe,
                       const double *ue, const double pt[3], double strain[6],
                       double stress[6]) {
  const int n = 3 * npe;
  double G[30], J[9], Ji[9], B[180], D[36];
  orc_basis_grad(npe, pt, G);
  orc_jacobian(npe, G, xe, J);
  orc_inv3x3(J, Ji);
  memset(B, 0, sizeof(double) * 6 * n);
  orc_bmatrix(npe, Ji, G, B);
  memset(strain, 0, sizeof(double) * 6);
  gemv_acc(6, n, B, ue, strain);
  orc_dmatrix(E, nu, D);
  memset(stress, 0, sizeof(double) * 6);
  gemv_acc(6, 6, D, strain, stress);
}

/* Dilatational wave speed and element characteristic length: the minimum
   altitude 3V / A_max of the corner tetrahedron divided by 2 (tet4) or 4.35
   (tet10).  Against the exact element limit 2/omega_max of K and the lumped
   mass, the bare altitude overshoots by up to 1.91x for tet4 (1.63x for the
   regular tetrahedron); the factors keep dt_e below the limit for every
   shape (checked on random elements in the tests). */
#define ORC_TET4_LENGTH_FACTOR 2.0
#define ORC_TET10_LENGTH_FACTOR 4.35

double orc_element_dt(int npe, double E, double nu, double rho,
                      const double *xe) {
  double c = sqrt(E * (1 - nu) / ((1 + nu) * (1 - 2 * nu)) / rho);
  const double *P[4] = {xe, xe + 3, xe + 6, xe + 9};
  double e1[3], e2[3], e3[3];
  for (int k = 0; k < 3; k++) {
    e1[k] = P[1][k] - P[0][k];
    e2[k] = P[2][k] - P[0][k];
    e3[k] = P[3][k] - P[0][k];
  }
  double vol6 = e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) -
                e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
  static const int F[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
  double amax2 = 0.0; /* (2A)^2 */
  for (int f = 0; f < 4; f++) {
    const double *a = P[F[f][0]], *b = P[F[f][1]], *d = P[F[f][2]];
    double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
    double w[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
    double cx = u[1] * w[2] - u[2] * w[1];
    double cy = u[2] * w[0] - u[0] * w[2];
    double cz = u[0] * w[1] - u[1] * w[0];
    double s = cx * cx + cy * cy + cz * cz;
    if (s > amax2) amax2 = s;
  }
  double h = fabs(vol6) / sqrt(amax2); /* 3V/A = (6V)/(2A) */
  h /= (npe == 10) ? ORC_TET10_LENGTH_FACTOR : ORC_TET4_LENGTH_FACTOR;
  return h / c;
}

/* ---------------------------------------------------------------------- */
/* J2 plasticity                                                           */
/* ---------------------------------------------------------------------- */

/* Return mapping for linear isotropic hardening, Voigt order xx yy zz xy yz
   xz with engineering shear strains (Simo & Hughes 1998, Box 3.2 restricted
   to isotropic hardening).  Yield: ||s|| <= sqrt(2/3) (sy0 + H ep). */
double orc_radial_return(double E, double nu, double sy0, double H,
                         const double de[6], double s[6], double *ep) {
  double G = E / (2.0 * (1.0 + nu));
  double lam = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  double tr = de[0] + de[1] + de[2];
  s[0] += lam * tr + 2.0 * G * de[0];
  s[1] += lam * tr + 2.0 * G * de[1];
  s[2] += lam * tr + 2.0 * G * de[2];
  s[3] += G * de[3];
  s[4] += G * de[4];
  s[5] += G * de[5];
  double p = (s[0] + s[1] + s[2]) / 3.0;
  double d0 = s[0] - p, d1 = s[1] - p, d2 = s[2] - p;
  double nrm = sqrt(d0 * d0 + d1 * d1 + d2 * d2 +
                    2.0 * (s[3] * s[3] + s[4] * s[4] + s[5] * s[5]));
  double fy = nrm - sqrt(2.0 / 3.0) * (sy0 + H * (*ep));
  if (fy <= 0.0) return 0.0;
  double dg = fy / (2.0 * G + (2.0 / 3.0) * H);
  double sc = 2.0 * G * dg / nrm;
  s[0] -= sc * d0;
  s[1] -= sc * d1;
  s[2] -= sc * d2;
  s[3] -= sc * s[3];
  s[4] -= sc * s[4];
  s[5] -= sc * s[5];
  *ep += sqrt(2.0 / 3.0) * dg;
  return dg;
}

void orc_element_force_ep(int npe, int rule, double E, double nu, double sy0,
                          double H, const double *xe, const double *due,
                          double *sig, double *ep, double *fe,
                          double work[2]) {
  double pt[3], G[30], J[9], Ji[9], dN[30];
  memset(fe, 0, sizeof(double) * 3 * npe);
  for (int q = 0; q < orc_quad_npts(rule); q++) {
    double w = orc_quad_point(rule, q, pt);
    orc_basis_grad(npe, pt, G);
    orc_jacobian(npe, G, xe, J);
    double detJ = orc_inv3x3(J, Ji);
    dndx_of(npe, Ji, G, dN);
    double de[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < npe; i++) {
      const double *d = dN + 3 * i, *u = due + 3 * i;
      de[0] += d[0] * u[0];
      de[1] += d[1] * u[1];
      de[2] += d[2] * u[2];
      de[3] += d[1] * u[0] + d[0] * u[1];
      de[4] += d[2] * u[1] + d[1] * u[2];
      de[5] += d[2] * u[0] + d[0] * u[2];
    }
    double *s = sig + 6 * q;
    double old[6];
    memcpy(old, s, sizeof(old));
    double dg = orc_radial_return(E, nu, sy0, H, de, s, ep + q);
    double wd = w * detJ;
    double dw = 0.0;
    for (int k = 0; k < 6; k++) dw += 0.5 * (old[k] + s[k]) * de[k];
    work[0] += wd * dw;
    if (dg > 0.0) work[1] += wd * sqrt(2.0 / 3.0) * (sy0 + H * ep[q]) * dg;
    for (int i = 0; i < npe; i++) {
      const double *d = dN + 3 * i;
      fe[3 * i + 0] += wd * (d[0] * s[0] + d[1] * s[3] + d[2] * s[5]);
      fe[3 * i + 1] += wd * (d[1] * s[1] + d[0] * s[3] + d[2] * s[4]);
      fe[3 * i + 2] += wd * (d[2] * s[2] + d[1] * s[4] + d[0] * s[5]);
    }
  }
}

/* ---------------------------------------------------------------------- */
/* known-answer driver                                                     */
/* ---------------------------------------------------------------------- */

/* dynamics.h:303-412: du = a x (1, -nu, -nu) (case 0) or a x^2/2 (1, -nu,
   -nu) (case 1); sum of element strain energies and volumes. */
void orc_debug_strain(int npe, int rule, double E, double nu, int nn, int ne,
                      const int *conn, const double *X, double al, int dc,
                      double out[2]) {
  double *du = (double *)calloc((size_t)3 * nn, sizeof(double));
  for (int i = 0; i < nn; i++) {
    double x = X[3 * i];
    if (dc == 0) {
      du[3 * i + 0] = al * x;
      du[3 * i + 1] = -al * x * nu;
      du[3 * i + 2] = -al * x * nu;
    } else if (dc == 1) {
      du[3 * i + 0] = al * 0.5 * x * x;
      du[3 * i + 1] = -al * 0.5 * x * x * nu;
      du[3 * i + 2] = -al * 0.5 * x * x * nu;
    }
  }
  double W = 0.0, V = 0.0, xe[30], ue[30];
  for (int e = 0; e < ne; e++) {
    for (int j = 0; j < npe; j++)
      for (int k = 0; k < 3; k++) {
        xe[3 * j + k] = X[3 * conn[npe * e + j] + k];
        ue[3 * j + k] = du[3 * conn[npe * e + j] + k];
      }
    W += orc_strain_energy(npe, rule, E, nu, xe, ue);
    V += orc_element_volume(npe, rule, xe);
  }
  out[0] = W;
  out[1] = V;
  free(du);
}

/* ---------------------------------------------------------------------- */
/* greedy colouring                                                        */
/* ---------------------------------------------------------------------- */

int orc_greedy_coloring(int ne, int npe, const int *conn, int nn, int *color) {
  int words = 1;
  uint64_t *mask = (uint64_t *)calloc((size_t)nn * words, sizeof(uint64_t));
  if (!mask) return -1;
  int ncol = 0;
  for (int e = 0; e < ne; e++) {
    const int *c = conn + (size_t)npe * e;
    int found = -1;
    for (;;) {
      for (int w = 0; w < words && found < 0; w++) {
        uint64_t used = 0;
        for (int j = 0; j < npe; j++) used |= mask[(size_t)c[j] * words + w];
        for (int b = 0; b < 64; b++)
          if (!((used >> b) & 1u)) { found = 64 * w + b; break; }
      }
      if (found >= 0) break;
      /* grow the per-node colour sets by one word */
      uint64_t *m2 = (uint64_t *)calloc((size_t)nn * (words + 1), sizeof(uint64_t));
      if (!m2) { free(mask); return -1; }
      for (int i = 0; i < nn; i++)
        memcpy(m2 + (size_t)i * (words + 1), mask + (size_t)i * words,
               sizeof(uint64_t) * words);
      free(mask);
      mask = m2;
      words++;
    }
    color[e] = found;
    if (found + 1 > ncol) ncol = found + 1;
    for (int j = 0; j < npe; j++)
      mask[(size_t)c[j] * words + found / 64] |= (uint64_t)1 << (found % 64);
  }
  free(mask);
  return ncol;
}

/* ---------------------------------------------------------------------- */
/* global step                                                             */
/* ---------------------------------------------------------------------- */

static void gather(int npe, const int *c, const double *g, double *e) {
  for (int j = 0; j < npe; j++)
    for (int k = 0; k < 3; k++) e[3 * j + k] = g[3 * c[j] + k];
}

double orc_stable_dt(const orc_problem *p, double *dt_contact) {
  double dte = INFINITY, xe[30];
  double mmin = INFINITY;
  double *m = (double *)calloc((size_t)p->num_nodes, sizeof(double));
  for (int e = 0; e < p->num_elements; e++) {
    const int *c = p->conn + (size_t)p->npe * e;
    gather(p->npe, c, p->X0, xe);
    double d = orc_element_dt(p->npe, p->E, p->nu, p->rho, xe);
    if (d < dte) dte = d;
    double me[10];
    orc_element_mass(p->npe, p->quad_rule, p->mass_kind, p->rho, xe, me);
    for (int j = 0; j < p->npe; j++) m[c[j]] += me[j];
  }
  for (int i = 0; i < p->num_nodes; i++)
    if (m[i] > 0.0 && m[i] < mmin) mmin = m[i];
  free(m);
  if (dt_contact) {
    *dt_contact = (p->wall_enabled && p->wall_stiffness > 0.0)
                      ? 2.0 * sqrt(mmin / p->wall_stiffness)
                      : INFINITY;
  }
  return dte;
}

/* The node loop of update() for node i (dynamics_kernels.h:150-168): wall
   penalty at x + du, then gravity; a held node keeps zero acceleration. */
static inline void node_external(const orc_problem *p, orc_state *s,
                                 const double *du, int i) {
  const int dim = p->wall_dim;
  const double nk = p->wall_normal * p->wall_stiffness;
  const int held = p->fixed_mask && p->fixed_mask[i];
  if (p->wall_enabled && !held && (!p->contact_mask || p->contact_mask[i])) {
    double pos = s->x[3 * i + dim] + du[3 * i + dim];
    double dist = (pos - p->wall_location) * p->wall_normal;
    if (dist < 0.0) s->a[3 * i + dim] += -1 * (1 / s->m[i]) * dist * nk;
  }
  s->a[3 * i + 0] += p->gravity[0];
  s->a[3 * i + 1] += p->gravity[1];
  s->a[3 * i + 2] += p->gravity[2];
  if (held) s->a[3 * i + 0] = s->a[3 * i + 1] = s->a[3 * i + 2] = 0.0;
}

/* One update() call (dynamics_kernels.h:8-169 for compat): zero the
   acceleration, assemble the mass on the first call, element loop with
   per-element -f_e/m accumulation, then the node loop with the wall penalty
   evaluated at x + du followed by gravity. */
static void update(const orc_problem *p, orc_state *s, const double *du,
                   double *wint, double *wpl, double *mindt) {
  const int nn = p->num_nodes, ne = p->num_elements, npe = p->npe;
  const int nd = 3 * npe;
  const int nq = orc_quad_npts(p->quad_rule);
  double xe[30], due[30], fe[30], me[10], xr[30], tot[30];
  memset(s->a, 0, sizeof(double) * 3 * nn);
  if (s->m[0] == 0.0) {
    for (int e = 0; e < ne; e++) {
      const int *c = p->conn + (size_t)npe * e;
      gather(npe, c, s->x, xe);
      orc_element_mass(npe, p->quad_rule, p->mass_kind, p->rho, xe, me);
      for (int j = 0; j < npe; j++) s->m[c[j]] += me[j];
    }
  }
  double md = INFINITY;
  for (int e = 0; e < ne; e++) {
    const int *c = p->conn + (size_t)npe * e;
    gather(npe, c, s->x, xe);
    gather(npe, c, du, due);
    if (p->physics == ORC_COMPAT) {
      double minv[30];
      for (int j = 0; j < npe; j++)
        for (int k = 0; k < 3; k++) minv[3 * j + k] = 1.0 / s->m[c[j]];
      orc_element_force_compat(npe, p->quad_rule, p->E, p->nu, xe, due, fe);
      double dw = 0.0;
      for (int j = 0; j < nd; j++) dw += fe[j] * due[j];
      *wint += dw;
      for (int j = 0; j < nd; j++) fe[j] = minv[j] * (-fe[j]);
      if (s->strain && s->stress) {
        double st[6], ss[6];
        const double pt0[3] = {0.0, 0.0, 0.0};
        gather(npe, c, s->Xref, xr);
        for (int j = 0; j < nd; j++) tot[j] = due[j] + xe[j] - xr[j];
        memset(st, 0, sizeof(st));
        memset(ss, 0, sizeof(ss));
        orc_stress_strain(npe, p->E, p->nu, xe, tot, pt0, st, ss);
        for (int j = 0; j < npe; j++) s->a[3 * c[j] + 0] += fe[3 * j + 0],
                                       s->a[3 * c[j] + 1] += fe[3 * j + 1],
                                       s->a[3 * c[j] + 2] += fe[3 * j + 2];
        for (int j = 0; j < npe; j++)
          for (int k = 0; k < 6; k++) {
            s->strain[6 * c[j] + k] = st[k];
            s->stress[6 * c[j] + k] = ss[k];
          }
      } else {
        for (int j = 0; j < npe; j++)
          for (int k = 0; k < 3; k++) s->a[3 * c[j] + k] += fe[3 * j + k];
      }
    } else {
      double w2[2] = {0.0, 0.0};
      double d = orc_element_dt(npe, p->E, p->nu, p->rho, xe);
      if (d < md) md = d;
      orc_element_force_ep(npe, p->quad_rule, p->E, p->nu, p->sigma_y, p->H,
                           xe, due, s->sig + (size_t)6 * nq * e,
                           s->epbar + (size_t)nq * e, fe, w2);
      *wint += w2[0];
      *wpl += w2[1];
      for (int j = 0; j < npe; j++)
        for (int k = 0; k < 3; k++) s->a[3 * c[j] + k] += fe[3 * j + k];
    }
  }
  if (p->physics != ORC_COMPAT) {
    for (int i = 0; i < nn; i++)
      for (int k = 0; k < 3; k++) s->a[3 * i + k] = -s->a[3 * i + k] / s->m[i];
  }
  *mindt = md;
  for (int i = 0; i < nn; i++) node_external(p, s, du, i);
}

static void energy_row(const orc_problem *p, const orc_state *s,
                       const double *vel, double *row) {
  double ke = 0.0, pen = 0.0, fc = 0.0;
  const double k = p->wall_stiffness;
  for (int i = 0; i < p->num_nodes; i++) {
    const double *v = vel + 3 * i;
    ke += 0.5 * s->m[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    const int held = p->fixed_mask && p->fixed_mask[i];
    if (p->wall_enabled && !held && (!p->contact_mask || p->contact_mask[i])) {
      double d = (s->x[3 * i + p->wall_dim] - p->wall_location) * p->wall_normal;
      if (d < 0.0) {
        pen += 0.5 * k * d * d;
        fc += -k * d;
      }
    }
  }
  row[ORC_E_KINETIC] = ke;
  row[ORC_E_PENALTY] = pen;
  row[ORC_E_CONTACTF] = fc;
}

static double now_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---------------------------------------------------------------------- */
/* threaded step (same arithmetic, same summation order)                   */
/* ---------------------------------------------------------------------- */

/* The threaded update splits the serial one in two passes: each element's
   nodal contributions and work go to per-element buffers (any order), then
   each node sums its contributions in ascending element order, the order in
   which the serial element loop adds them.  The work terms are summed in
   element order on one thread.  Every floating-point operation is the
   serial one, so the results are bit-identical to orc_solve. */
#define ORC_MAX_THREADS 64

typedef struct par_ctx par_ctx;
typedef void (*range_fn)(par_ctx *, int lo, int hi);
struct par_ctx {
  const orc_problem *p;
  orc_state *s;
  const double *du;
  double *fbuf;      /* [E][npe][3] -f_e/m (compat) or f_e (ep) */
  double *wbuf;      /* [E][2] internal / plastic work of the element */
  double *dtbuf;     /* [E] element stable dt (ep) */
  int *nstart;       /* [N+1] node -> first slot */
  int *nslot;        /* [E*npe] slots e*npe+j in ascending e */
  double dt;
  int nthreads;
  range_fn fn;
};

typedef struct {
  par_ctx *c;
  int lo, hi;
} par_job;

static void *par_run(void *a) {
  par_job *j = (par_job *)a;
  j->c->fn(j->c, j->lo, j->hi);
  return NULL;
}

/* fn over [0, n) in nthreads static chunks; chunk 0 (and any chunk whose
   thread cannot be started) runs on the calling thread */
static void par_for(par_ctx *c, range_fn fn, int n) {
  const int nt = c->nthreads;
  pthread_t th[ORC_MAX_THREADS];
  par_job jb[ORC_MAX_THREADS];
  int started[ORC_MAX_THREADS];
  c->fn = fn;
  for (int t = 0; t < nt; t++) {
    jb[t].c = c;
    jb[t].lo = (int)((long)n * t / nt);
    jb[t].hi = (int)((long)n * (t + 1) / nt);
    started[t] = t > 0 && pthread_create(&th[t], NULL, par_run, &jb[t]) == 0;
  }
  for (int t = 0; t < nt; t++)
    if (!started[t]) par_run(&jb[t]);
  for (int t = 1; t < nt; t++)
    if (started[t]) pthread_join(th[t], NULL);
}

/* the element loop of update() for elements [lo, hi) */
static void par_elements(par_ctx *c, int lo, int hi) {
  const orc_problem *p = c->p;
  const orc_state *s = c->s;
  const int npe = p->npe, nd = 3 * npe;
  const int nq = orc_quad_npts(p->quad_rule);
  double xe[30], due[30];
  for (int e = lo; e < hi; e++) {
    const int *cn = p->conn + (size_t)npe * e;
    double *fe = c->fbuf + (size_t)nd * e;
    gather(npe, cn, s->x, xe);
    gather(npe, cn, c->du, due);
    if (p->physics == ORC_COMPAT) {
      double minv[30];
      for (int j = 0; j < npe; j++)
        for (int k = 0; k < 3; k++) minv[3 * j + k] = 1.0 / s->m[cn[j]];
      orc_element_force_compat(npe, p->quad_rule, p->E, p->nu, xe, due, fe);
      double dw = 0.0;
      for (int j = 0; j < nd; j++) dw += fe[j] * due[j];
      c->wbuf[2 * (size_t)e] = dw;
      c->wbuf[2 * (size_t)e + 1] = 0.0;
      for (int j = 0; j < nd; j++) fe[j] = minv[j] * (-fe[j]);
    } else {
      double w2[2] = {0.0, 0.0};
      c->dtbuf[e] = orc_element_dt(npe, p->E, p->nu, p->rho, xe);
      orc_element_force_ep(npe, p->quad_rule, p->E, p->nu, p->sigma_y, p->H,
                           xe, due, s->sig + (size_t)6 * nq * e,
                           s->epbar + (size_t)nq * e, fe, w2);
      c->wbuf[2 * (size_t)e] = w2[0];
      c->wbuf[2 * (size_t)e + 1] = w2[1];
    }
  }
}

/* the node side of update() for nodes [lo, hi) */
static void par_nodes(par_ctx *c, int lo, int hi) {
  const orc_problem *p = c->p;
  orc_state *s = c->s;
  for (int i = lo; i < hi; i++) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int q = c->nstart[i]; q < c->nstart[i + 1]; q++) {
      const double *f = c->fbuf + 3 * (size_t)c->nslot[q];
      a0 += f[0];
      a1 += f[1];
      a2 += f[2];
    }
    s->a[3 * i + 0] = a0;
    s->a[3 * i + 1] = a1;
    s->a[3 * i + 2] = a2;
    if (p->physics != ORC_COMPAT)
      for (int k = 0; k < 3; k++) s->a[3 * i + k] = -s->a[3 * i + k] / s->m[i];
    node_external(p, s, c->du, i);
  }
}

/* the loop body's nodal update (x, v, v_i) for dofs [lo, hi) */
static void par_kick(par_ctx *c, int lo, int hi) {
  orc_state *s = c->s;
  const double *du = c->du;
  const double dt = c->dt;
  for (int i = lo; i < hi; i++) {
    s->x[i] += du[i];
    s->v[i] += dt * s->a[i];
    s->v_i[i] = s->v[i] - 0.5 * dt * s->a[i];
  }
}

static void update_par(par_ctx *c, const double *du, double *wint, double *wpl,
                       double *mindt) {
  const orc_problem *p = c->p;
  const int ne = p->num_elements;
  c->du = du;
  par_for(c, par_elements, ne);
  double md = INFINITY;
  for (int e = 0; e < ne; e++) {
    *wint += c->wbuf[2 * (size_t)e];
    if (p->physics != ORC_COMPAT) {
      *wpl += c->wbuf[2 * (size_t)e + 1];
      if (c->dtbuf[e] < md) md = c->dtbuf[e];
    }
  }
  *mindt = md;
  par_for(c, par_nodes, p->num_nodes);
}

static void par_free(par_ctx *c) {
  free(c->fbuf);
  free(c->wbuf);
  free(c->dtbuf);
  free(c->nstart);
  free(c->nslot);
}

/* buffers and the node -> slot lists; 0 on success */
static int par_setup(par_ctx *c, const orc_problem *p, orc_state *s,
                     int nthreads) {
  const int nn = p->num_nodes, ne = p->num_elements, npe = p->npe;
  const size_t ns = (size_t)ne * npe;
  memset(c, 0, sizeof(*c));
  c->p = p;
  c->s = s;
  c->dt = p->dt;
  c->nthreads = nthreads > ORC_MAX_THREADS ? ORC_MAX_THREADS : nthreads;
  c->fbuf = (double *)malloc(sizeof(double) * 3 * (ns ? ns : 1));
  c->wbuf = (double *)malloc(sizeof(double) * 2 * (ne ? ne : 1));
  c->dtbuf = (double *)malloc(sizeof(double) * (ne ? ne : 1));
  c->nstart = (int *)calloc((size_t)nn + 1, sizeof(int));
  c->nslot = (int *)malloc(sizeof(int) * (ns ? ns : 1));
  if (!c->fbuf || !c->wbuf || !c->dtbuf || !c->nstart || !c->nslot) {
    par_free(c);
    return -1;
  }
  for (size_t q = 0; q < ns; q++) c->nstart[p->conn[q] + 1]++;
  for (int i = 0; i < nn; i++) c->nstart[i + 1] += c->nstart[i];
  int *fill = (int *)malloc(sizeof(int) * (nn ? nn : 1));
  if (!fill) {
    par_free(c);
    return -1;
  }
  memcpy(fill, c->nstart, sizeof(int) * nn);
  for (size_t q = 0; q < ns; q++) c->nslot[fill[p->conn[q]]++] = (int)q;
  free(fill);
  return 0;
}

static int solve_impl(const orc_problem *p, int nsteps, orc_state *s,
                      double *loop_seconds, par_ctx *pc);

int orc_solve(const orc_problem *p, int nsteps, orc_state *s) {
  return solve_impl(p, nsteps, s, NULL, NULL);
}

int orc_solve_timed(const orc_problem *p, int nsteps, orc_state *s,
                    double *loop_seconds) {
  return solve_impl(p, nsteps, s, loop_seconds, NULL);
}

int orc_solve_threaded(const orc_problem *p, int nsteps, orc_state *s,
                       int nthreads, double *loop_seconds) {
  /* the last-writer-wins strain/stress output follows the serial loop */
  if (nthreads <= 1 || (s->strain && s->stress))
    return solve_impl(p, nsteps, s, loop_seconds, NULL);
  par_ctx c;
  if (par_setup(&c, p, s, nthreads) != 0) return -1;
  int rc = solve_impl(p, nsteps, s, loop_seconds, &c);
  par_free(&c);
  return rc;
}

static int solve_impl(const orc_problem *p, int nsteps, orc_state *s,
                      double *loop_seconds, par_ctx *pc) {
  const int nn = p->num_nodes, nd = 3 * nn;
  const int nq = orc_quad_npts(p->quad_rule);
  const double dt = p->dt;
  double *du = (double *)calloc((size_t)nd, sizeof(double));
  if (!du) return -1;
  memcpy(s->x, p->X0, sizeof(double) * nd);
  memcpy(s->Xref, p->X0, sizeof(double) * nd);
  memcpy(s->v, p->v0, sizeof(double) * nd);
  if (p->fixed_mask)
    for (int i = 0; i < nn; i++)
      if (p->fixed_mask[i]) s->v[3 * i] = s->v[3 * i + 1] = s->v[3 * i + 2] = 0.0;
  memset(s->a, 0, sizeof(double) * nd);
  memset(s->m, 0, sizeof(double) * nn);
  memset(s->v_i, 0, sizeof(double) * nd);
  if (p->physics != ORC_COMPAT) {
    memset(s->sig, 0, sizeof(double) * 6 * nq * (size_t)p->num_elements);
    memset(s->epbar, 0, sizeof(double) * nq * (size_t)p->num_elements);
  }
  double wint = 0.0, wpl = 0.0, wgrav = 0.0, mindt = 0.0;
  update(p, s, du, &wint, &wpl, &mindt);
  double e0 = 0.0;
  if (s->energy) {
    double *r = s->energy;
    energy_row(p, s, s->v, r);
    r[ORC_E_INTERNAL] = wint;
    r[ORC_E_PLASTIC] = wpl;
    r[ORC_E_GRAVITY] = 0.0;
    r[ORC_E_BALANCE] = 0.0;
    r[ORC_E_MINDT] = mindt;
    e0 = r[ORC_E_KINETIC] + r[ORC_E_PENALTY];
  }
  for (int i = 0; i < nd; i++) s->v[i] += 0.5 * dt * s->a[i];
  double t0 = now_seconds();
  for (int n = 1; n <= nsteps; n++) {
    for (int j = 0; j < nd; j++) du[j] = dt * s->v[j];
    if (pc) {
      update_par(pc, du, &wint, &wpl, &mindt);
      par_for(pc, par_kick, nd);
    } else {
      update(p, s, du, &wint, &wpl, &mindt);
      for (int i = 0; i < nd; i++) {
        s->x[i] += du[i];
        s->v[i] += dt * s->a[i];
        s->v_i[i] = s->v[i] - 0.5 * dt * s->a[i];
      }
    }
    if (s->energy) {
      for (int i = 0; i < nn; i++)
        for (int k = 0; k < 3; k++)
          wgrav += s->m[i] * p->gravity[k] * du[3 * i + k];
      double *r = s->energy + (size_t)ORC_NENERGY * n;
      energy_row(p, s, s->v_i, r);
      r[ORC_E_INTERNAL] = wint;
      r[ORC_E_PLASTIC] = wpl;
      r[ORC_E_GRAVITY] = wgrav;
      r[ORC_E_BALANCE] = r[ORC_E_KINETIC] + wint + r[ORC_E_PENALTY] - wgrav - e0;
      r[ORC_E_MINDT] = mindt;
    }
  }
  if (loop_seconds) *loop_seconds = now_seconds() - t0;
  free(du);
  return 0;
}
