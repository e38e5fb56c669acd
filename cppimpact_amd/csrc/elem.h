// Element physics on the device: tet4 and tet10 internal force, stress
// update, lumped mass and the stable-dt estimate.  Included by kernels.hip.
//
// The element cores read nodes through a getter get(k, x[3], v[3]) and hand
// corner forces to a putter put(k, f[3]), so the same physics serves the
// LDS-staged workgroup kernels and the one-element-per-thread kernels
// without holding every node record in registers.
#pragma once

#include "device.h"
#include "mlp.h"

namespace cpi {

enum : int { PHYS_COMPAT = 0, PHYS_EP = 1, PHYS_EP_MLP = 2 };
enum : int { RULE_REF5 = 0, RULE_REF4 = 1, RULE_EXACT4 = 2, RULE_CENTROID = 3 };

// Quadrature rules as compile-time constants so that the tet10 basis
// gradients at the points fold into immediates.
template <int RULE> struct Quad;
template <> struct Quad<RULE_REF5> {  // tetrahedral.h:12-40
  static constexpr int n = 5;
  __device__ static constexpr double w(int q) { return q == 0 ? -0.13333333333 : 0.075; }
  __device__ static constexpr double pt(int q, int c) {
    return q == 0 ? 0.25 : (q - 1 == c ? 0.5 : 0.16666666666);
  }
};
template <> struct Quad<RULE_REF4> {  // tetrahedral.h:48-73
  static constexpr int n = 4;
  __device__ static constexpr double w(int) { return 0.041666667; }
  __device__ static constexpr double pt(int q, int c) {
    return (q >= 1 && q - 1 == c) ? 0.58541020 : 0.13819660;
  }
};
template <> struct Quad<RULE_EXACT4> {  // alpha = (5+3 sqrt5)/20, beta = (5-sqrt5)/20
  static constexpr int n = 4;
  __device__ static constexpr double w(int) { return 1.0 / 24.0; }
  __device__ static constexpr double pt(int q, int c) {
    return (q >= 1 && q - 1 == c) ? 0.58541019662496845446 : 0.13819660112501051518;
  }
};
template <> struct Quad<RULE_CENTROID> {
  static constexpr int n = 1;
  __device__ static constexpr double w(int) { return 1.0 / 6.0; }
  __device__ static constexpr double pt(int, int) { return 0.25; }
};

// Stress-state words are read and written once per step.  STREAM = true
// marks those accesses non-temporal: for a state larger than the
// last-level cache, caching it only evicts the node records and partial
// forces that other blocks and the boundary-node kernel read again.
template <bool STREAM, typename T>
__device__ __forceinline__ T state_ld(const T *a) {
  if constexpr (STREAM)
    return __builtin_nontemporal_load(a);
  else
    return *a;
}
template <bool STREAM, typename T>
__device__ __forceinline__ void state_st(T *a, T v) {
  if constexpr (STREAM)
    __builtin_nontemporal_store(v, a);
  else
    *a = v;
}

template <typename R, bool STREAM = false>
__device__ __forceinline__ void load_state(const Params &p, int e, int q, R s[6], R &ep) {
  const R *S = (const R *)p.sig;
  const R *P = (const R *)p.epbar;
  const long ne = p.ne;
#pragma unroll
  for (int c = 0; c < 6; c++) s[c] = state_ld<STREAM>(&S[(long)(q * 6 + c) * ne + e]);
  ep = state_ld<STREAM>(&P[(long)q * ne + e]);
}

// ep is written only when it changed (a plastic step): elastic steps store
// the stress alone
template <typename R, bool STREAM = false>
__device__ __forceinline__ void store_state(const Params &p, int e, int q, const R s[6], R ep,
                                            bool ep_changed = true) {
  R *S = (R *)p.sig;
  R *P = (R *)p.epbar;
  const long ne = p.ne;
#pragma unroll
  for (int c = 0; c < 6; c++) state_st<STREAM>(&S[(long)(q * 6 + c) * ne + e], s[c]);
  if (ep_changed) state_st<STREAM>(&P[(long)q * ne + e], ep);
}

// State of one point loaded ahead of use (the tet4 workgroup kernel issues
// it before its first barrier so the load overlaps the node staging);
// point_stress leaves the updated ep in it, and with DEFER the updated
// stress too, for the caller to store (changed: ep changed).
template <typename R> struct PreState {
  R s[6];
  R ep;
  bool changed;
};

// The stress of one point, and its ep only when the point has yielded
// before (ep of the others is 0).
template <typename R, bool STREAM = false>
__device__ __forceinline__ void load_state_if(const Params &p, int e, int q, bool has_ep,
                                              PreState<R> &st) {
  const R *S = (const R *)p.sig;
  const long ne = p.ne;
#pragma unroll
  for (int c = 0; c < 6; c++) st.s[c] = state_ld<STREAM>(&S[(long)(q * 6 + c) * ne + e]);
  st.ep = has_ep ? state_ld<STREAM>(&((const R *)p.epbar)[(long)q * ne + e]) : R(0);
}

// A zero word in global memory (the bits of 0.0 in FP64 and in FP32); no
// kernel writes it.  Not const: a constant-memory address selected against
// a global one makes the load a flat load, which the compiler waits for in
// full before the next barrier.
__device__ double kZeroWord[1] = {0.0};

// load_state_if without a branch, for the block kernels' load phase: a
// point that has not yielded loads its ep from kZeroWord, so every lane
// issues the same seven loads, nothing waits for them before their first
// use, and the kernel can wait for the loads issued before them by count
// while these stay in flight.
template <typename R, bool STREAM = false>
__device__ __forceinline__ void load_state_sel(const Params &p, int e, int q, bool has_ep,
                                               PreState<R> &st) {
  const R *S = (const R *)p.sig;
  const R *P = (const R *)p.epbar;
  const long ne = p.ne;
#pragma unroll
  for (int c = 0; c < 6; c++) st.s[c] = state_ld<STREAM>(&S[(long)(q * 6 + c) * ne + e]);
  st.ep = state_ld<STREAM>(has_ep ? &P[(long)q * ne + e] : (const R *)kZeroWord);
}

// Stress at one point from the strain increment de.  compat: sigma = D de
// with no memory (dynamics_kernels.h:101-114 evaluates K_e du); ep: trial
// sigma_n + D de, then the return map (exact or surrogate).  Adds the
// point's internal work and plastic dissipation (times wd) to ea.  (The
// surrogate's return map is applied by the tet4 block kernel's walk.)
template <typename R, int PHYS, bool PRE = false, bool STREAM = false, bool DEFER = false>
__device__ __forceinline__ void point_stress(const Params &p, int e, int q, const R de[6],
                                             R wd, R s[6], EAcc &ea,
                                             PreState<R> *pre = nullptr) {
  const R c11 = R(p.c11), c12 = R(p.c12), c44 = R(p.c44);
  R ds[6];
  ds[0] = c11 * de[0] + c12 * (de[1] + de[2]);
  ds[1] = c11 * de[1] + c12 * (de[0] + de[2]);
  ds[2] = c11 * de[2] + c12 * (de[0] + de[1]);
  ds[3] = c44 * de[3];
  ds[4] = c44 * de[4];
  ds[5] = c44 * de[5];
  if constexpr (PHYS == PHYS_COMPAT) {
    R w = R(0);
#pragma unroll
    for (int c = 0; c < 6; c++) {
      s[c] = ds[c];
      w += s[c] * de[c];
    }
    ea.wint += double(wd * w);
  } else {
    R so[6], ep;
    if constexpr (PRE) {
#pragma unroll
      for (int c = 0; c < 6; c++) so[c] = pre->s[c];
      ep = pre->ep;
    } else {
      load_state<R>(p, e, q, so, ep);
    }
    // midpoint internal work 1/2 (sigma_n + sigma_{n+1}) : de, split so that
    // sigma_n need not stay live through the return map
    R w = R(0);
#pragma unroll
    for (int c = 0; c < 6; c++) {
      w += so[c] * de[c];
      s[c] = so[c] + ds[c];
    }
    const R ep0 = ep;
    const R dg = radial_return<R>(p, s, ep);
    if (p.samples && dg > R(0)) record_sample<R>(p, e, q, de, so, ep0, s, ep);
#pragma unroll
    for (int c = 0; c < 6; c++) w += s[c] * de[c];
    ea.wint += double(wd * R(0.5) * w);
    if (dg > R(0))
      ea.wpl += double(wd * R(0.81649658092772603273) * (R(p.sy0) + R(p.H) * ep) * dg);
    if constexpr (PRE && DEFER) {
#pragma unroll
      for (int c = 0; c < 6; c++) pre->s[c] = s[c];
      pre->changed = dg > R(0);
    } else {
      store_state<R, STREAM>(p, e, q, s, ep, dg > R(0));
    }
    if constexpr (PRE) pre->ep = ep;
  }
}

// trial stress s = sigma_n + D de; returns sigma_n : de (the first half of
// the midpoint work)
template <typename R>
__device__ __forceinline__ R ep_trial(const Params &p, const R de[6], const R so[6], R s[6]) {
  const R c11 = R(p.c11), c12 = R(p.c12), c44 = R(p.c44);
  R ds[6];
  ds[0] = c11 * de[0] + c12 * (de[1] + de[2]);
  ds[1] = c11 * de[1] + c12 * (de[0] + de[2]);
  ds[2] = c11 * de[2] + c12 * (de[0] + de[1]);
  ds[3] = c44 * de[3];
  ds[4] = c44 * de[4];
  ds[5] = c44 * de[5];
  R w = R(0);
#pragma unroll
  for (int c = 0; c < 6; c++) {
    w += so[c] * de[c];
    s[c] = so[c] + ds[c];
  }
  return w;
}

// sigma . dN_i (engineering-shear Voigt order xx, yy, zz, xy, yz, xz)
template <typename R>
__device__ __forceinline__ void sig_dn(const R dn[3], const R s[6], R g[3]) {
  g[0] = dn[0] * s[0] + dn[1] * s[3] + dn[2] * s[5];
  g[1] = dn[1] * s[1] + dn[0] * s[3] + dn[2] * s[4];
  g[2] = dn[2] * s[2] + dn[1] * s[4] + dn[0] * s[5];
}

// f_i += wd * sigma . dN_i, one rounding per component (fused)
template <typename R>
__device__ __forceinline__ void add_bts(R f[3], const R dn[3], const R s[6], R wd) {
  R g[3];
  sig_dn<R>(dn, s, g);
#pragma unroll
  for (int c = 0; c < 3; c++) f[c] = fma(wd, g[c], f[c]);
}

// minimum altitude of the corner tetrahedron / wave speed
template <typename R>
__device__ __forceinline__ double corner_dt(const R e1[3], const R e2[3], const R e3[3], R vol6,
                                            double c, double factor) {
  auto cr2 = [](const R a[3], const R b[3]) {
    R x = a[1] * b[2] - a[2] * b[1], y = a[2] * b[0] - a[0] * b[2], z = a[0] * b[1] - a[1] * b[0];
    return x * x + y * y + z * z;
  };
  R d21[3] = {e2[0] - e1[0], e2[1] - e1[1], e2[2] - e1[2]};
  R d31[3] = {e3[0] - e1[0], e3[1] - e1[1], e3[2] - e1[2]};
  R a = cr2(d21, d31);
  a = max(a, cr2(e2, e3));
  a = max(a, cr2(e1, e3));
  a = max(a, cr2(e1, e2));
  return double(fabs(vol6)) / sqrt(double(a)) / (c * factor);
}

// characteristic length = minimum corner altitude / factor; the bare
// altitude overshoots the exact element limit 2/omega_max (stiffness against
// lumped mass) by up to 1.91x for tet4, hence 2; 4.35 keeps tet10 below it
constexpr double kTet4LengthFactor = 2.0;
constexpr double kTet10LengthFactor = 4.35;

// Kinematics of one tet4: inverse Jacobian, strain increment de (engineering
// shear) and w detJ; counts a non-positive detJ and, for the elastoplastic
// models, the element's stable dt into ea.
template <typename R, int PHYS, class Get>
__device__ __forceinline__ void tet4_kin(const Params &p, const Get &get, double dt, R Ji[9],
                                         R de[6], R &wd, EAcc &ea) {
  R J[9];
  R u[3][3];
  {
    double x0[3], v0[3];
    get(0, x0, v0);
#pragma unroll
    for (int k = 1; k < 4; k++) {
      double xk[3], vk[3];
      get(k, xk, vk);
#pragma unroll
      for (int a = 0; a < 3; a++) {
        J[a * 3 + (k - 1)] = R(xk[a] - x0[a]);
        u[k - 1][a] = R(dt * vk[a] - dt * v0[a]);
      }
    }
  }
  R detJ = inv3<R>(J, Ji);
  if (!(detJ > R(0))) ea.bad += 1.0;
  if constexpr (PHYS != PHYS_COMPAT) {
    R e1[3] = {J[0], J[3], J[6]}, e2[3] = {J[1], J[4], J[7]}, e3[3] = {J[2], J[5], J[8]};
    double d = corner_dt<R>(e1, e2, e3, detJ, p.wave_speed, kTet4LengthFactor);
    ea.mindt = fmin(ea.mindt, d);
  }
  // L[i][j] = d du_i / d x_j
  R L[3][3];
#pragma unroll
  for (int i = 0; i < 3; i++)
#pragma unroll
    for (int j = 0; j < 3; j++)
      L[i][j] = Ji[0 * 3 + j] * u[0][i] + Ji[1 * 3 + j] * u[1][i] + Ji[2 * 3 + j] * u[2][i];
  de[0] = L[0][0];
  de[1] = L[1][1];
  de[2] = L[2][2];
  de[3] = L[0][1] + L[1][0];
  de[4] = L[1][2] + L[2][1];
  de[5] = L[0][2] + L[2][0];
  wd = R(p.wsum) * detJ;
}

// corner forces w detJ sigma . dN_k of one tet4, handed to put() as known
template <typename R, class Put>
__device__ __forceinline__ void tet4_forces(const R Ji[9], const R s[6], R wd, const Put &put) {
  R f0[3] = {R(0), R(0), R(0)};
#pragma unroll
  for (int k = 1; k < 4; k++) {
    R dn[3] = {Ji[(k - 1) * 3 + 0], Ji[(k - 1) * 3 + 1], Ji[(k - 1) * 3 + 2]};
    R fk[3] = {R(0), R(0), R(0)};
    add_bts<R>(fk, dn, s, wd);
#pragma unroll
    for (int c = 0; c < 3; c++) f0[c] -= fk[c];
    put(k, fk);
  }
  put(0, f0);
}

// ---------------------------------------------------------------------------
// tet4: B is constant, every quadrature point carries the same strain, so
// the rule enters only through its weight sum (compat) and the element keeps
// one stress point (ep).  Corner forces reach put() as soon as known.
// STREAM: the state is read and written non-temporal.
// ---------------------------------------------------------------------------
template <typename R, int PHYS, bool PRE, bool STREAM = false, class Get, class Put>
__device__ __forceinline__ void tet4_core(const Params &p, const Get &get, double dt, int e,
                                          const Put &put, EAcc &ea,
                                          PreState<R> *pre = nullptr) {
  R Ji[9], de[6], wd;
  tet4_kin<R, PHYS>(p, get, dt, Ji, de, wd, ea);
  R s[6];
  point_stress<R, PHYS, PRE, STREAM>(p, e, 0, de, wd, s, ea, pre);
  tet4_forces<R>(Ji, s, wd, put);
}

// Entries of tet10_grad that are zero at every point (G[3i+c] = dN_i/dxi_c).
// Products with them are skipped: the compiler may not fold x * 0 away
// (0 * inf is NaN), so they would cost instructions and registers.
__host__ __device__ constexpr bool tet10_nz(int k) {
  return !(k == 1 || k == 2 || k == 3 || k == 5 || k == 9 || k == 10 || k == 14 || k == 22 ||
           k == 24);
}

// J[a][c] += G[3i+c] (x_i - o)_a over the non-zero G of node i
template <typename R, typename G>
__device__ __forceinline__ void tet10_jac_add(R J[9], const G Gd[30], int i, const R d[3]) {
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int c = 0; c < 3; c++)
      if (tet10_nz(3 * i + c)) J[a * 3 + c] += R(Gd[3 * i + c]) * d[a];
}

// dN_i/dx = Ji^T G_i, summed left to right over the non-zero G of node i
template <typename R, typename G>
__device__ __forceinline__ void tet10_dn(const R Ji[9], const G Gd[30], int i, R dn[3]) {
#pragma unroll
  for (int c = 0; c < 3; c++) {
    R acc = R(0);
    bool first = true;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if (!tet10_nz(3 * i + j)) continue;
      const R t = Ji[j * 3 + c] * R(Gd[3 * i + j]);
      acc = first ? t : acc + t;
      first = false;
    }
    dn[c] = acc;
  }
}

// ---------------------------------------------------------------------------
// tet10: loop over the rule's points with coordinates relative to corner 2
// (the origin of the reference's quadratic parametrisation); nodes are
// re-read through get() at every pass instead of being held in registers.
// Corner forces accumulate through add(i, dN_i, sigma, w detJ), which adds
// w detJ sigma . dN_i to node i's force, one point at a time in rule order.
// ---------------------------------------------------------------------------
template <int RULE, typename R, int PHYS, class Get, class Add>
__device__ __forceinline__ void tet10_points(const Params &p, const Get &get, double dt, int e,
                                             const Add &add, EAcc &ea) {
  double o[3];
  {
    double ov[3];
    get(2, o, ov);
  }
  using Q = Quad<RULE>;
  // a rolled loop: unrolled, the compiler hoists every point's state loads
  // and the kernel no longer fits the register file
#pragma unroll 1
  for (int q = 0; q < Q::n; q++) {
    double Gd[30];
    tet10_grad(Q::pt(q, 0), Q::pt(q, 1), Q::pt(q, 2), Gd);
    R J[9] = {R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0)};
#pragma unroll
    for (int i = 0; i < 10; i++) {
      double xi[3], vi[3];
      get(i, xi, vi);
      const R d[3] = {R(xi[0] - o[0]), R(xi[1] - o[1]), R(xi[2] - o[2])};
      tet10_jac_add<R>(J, Gd, i, d);
    }
    R Ji[9];
    R detJ = inv3<R>(J, Ji);
    if (!(detJ > R(0))) ea.bad += 1.0;
    R de[6] = {R(0), R(0), R(0), R(0), R(0), R(0)};
#pragma unroll
    for (int i = 0; i < 10; i++) {
      double xi[3], vi[3];
      get(i, xi, vi);
      R u[3] = {R(dt * vi[0]), R(dt * vi[1]), R(dt * vi[2])};
      R dn[3];
      tet10_dn<R>(Ji, Gd, i, dn);
      de[0] += dn[0] * u[0];
      de[1] += dn[1] * u[1];
      de[2] += dn[2] * u[2];
      de[3] += dn[1] * u[0] + dn[0] * u[1];
      de[4] += dn[2] * u[1] + dn[1] * u[2];
      de[5] += dn[2] * u[0] + dn[0] * u[2];
    }
    R wd = R(Q::w(q)) * detJ;
    R s[6];
    point_stress<R, PHYS>(p, e, q, de, wd, s, ea);
#pragma unroll
    for (int i = 0; i < 10; i++) {
      R dn[3];
      tet10_dn<R>(Ji, Gd, i, dn);
      add(i, dn, s, wd);
    }
  }
  if constexpr (PHYS != PHYS_COMPAT) {
    // corners in the linear ordering: (2, 0, 1, 3) is an even permutation
    R e1[3], e2[3], e3[3];
    double xa[3], va[3];
    get(0, xa, va);
#pragma unroll
    for (int c = 0; c < 3; c++) e1[c] = R(xa[c] - o[c]);
    get(1, xa, va);
#pragma unroll
    for (int c = 0; c < 3; c++) e2[c] = R(xa[c] - o[c]);
    get(3, xa, va);
#pragma unroll
    for (int c = 0; c < 3; c++) e3[c] = R(xa[c] - o[c]);
    R v6 = e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
           e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
    ea.mindt = fmin(ea.mindt, corner_dt<R>(e1, e2, e3, v6, p.wave_speed, kTet10LengthFactor));
  }
}

// tet10 on a 4-point rule with one lane per point: lanes 4k..4k+3 hold the
// same element e, lane q = (lane & 3) point q.  All four lanes of a quad
// must be active.  pre: the point's stress state, loaded ahead (EP models).
//
// The basis gradients of tet10_grad at a point, with a = 4 xi, b = 4 eta,
// c = 4 zeta and l = 4 (1 - xi - eta - zeta), are
//   G_0 = (a-1, 0, 0)    G_1 = (0, b-1, 0)    G_2 = (1-l)(1, 1, 1)
//   G_3 = (0, 0, c-1)    G_4 = (b, a, 0)      G_5 = (-b, l-b, -b)
//   G_6 = (l-a, -a, -a)  G_7 = (c, 0, a)      G_8 = (0, c, b)
//   G_9 = (-c, -c, l-c)
// and the point's arithmetic is written on that structure:
//   J  = sum_i (x_i - x_2) (x) G_i          (node 2 is the origin)
//   Lv = sum_i v_i (x) G_i, grad v = Lv J^-1, de = dt sym(grad v)
//   M  = w detJ sigma J^-T, f_i = M G_i      (= w detJ sigma . dN_i)
// so a node's force is one or two products per component instead of the
// gradient dN_i = J^-T G_i and sigma . dN_i.  Each corner force is then
// summed over the quad's lanes as (q0 + q1) + (q2 + q3) (quad_sum, DPP
// inside the quad), so every lane holds corner i's element force and hands
// it to put(i, f).  DEFER: the updated state is left in pre (ep changed
// in pre.changed) for the caller to store after everything else it does,
// else it is stored here (STREAM: non-temporal): registers that a store
// reads must not be reused before the store has completed, so a store
// issued here makes the rest of the kernel wait for it.
template <int RULE, typename R, int PHYS, bool STREAM, bool DEFER, class Get, class Put>
__device__ __forceinline__ void tet10_point_lane(const Params &p, const Get &get, double dt,
                                                 int e, int q, PreState<R> &pre, const Put &put,
                                                 EAcc &ea) {
  using Q = Quad<RULE>;
  static_assert(Q::n == 4, "one lane per point of a 4-point rule");
  const double px = Q::pt(q, 0), py = Q::pt(q, 1), pz = Q::pt(q, 2);
  const R a = R(4.0 * px), b = R(4.0 * py), c = R(4.0 * pz);
  const R l = R(4.0 - 4.0 * px - 4.0 * py - 4.0 * pz);
  // the differences are formed where they are used, not held for the
  // whole point
  auto am1 = [&]() { return a - R(1); };
  auto bm1 = [&]() { return b - R(1); };
  auto cm1 = [&]() { return c - R(1); };
  auto l1 = [&]() { return R(1) - l; };
  auto la = [&]() { return l - a; };
  auto lb = [&]() { return l - b; };
  auto lc = [&]() { return l - c; };
  double o[3];
  {
    double ov[3];
    get(2, o, ov);
  }
  auto rel = [&](int i, R d[3]) {
    double xi[3], vi[3];
    get(i, xi, vi);
#pragma unroll
    for (int r = 0; r < 3; r++) d[r] = R(xi[r] - o[r]);
  };
  auto vel = [&](int i, R u[3]) {
    double xi[3], vi[3];
    get(i, xi, vi);
#pragma unroll
    for (int r = 0; r < 3; r++) u[r] = R(vi[r]);
  };
  // J[r][k] = d x_r / d xi_k; the stable dt of the corner tetrahedron
  // (nodes 0, 1, 3 about node 2) from the same differences
  R J[9] = {R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0)};
  {
    R d[3], e1[3], e2[3];
    rel(0, e1);
#pragma unroll
    for (int r = 0; r < 3; r++) J[3 * r] += am1() * e1[r];
    rel(1, e2);
#pragma unroll
    for (int r = 0; r < 3; r++) J[3 * r + 1] += bm1() * e2[r];
    rel(3, d);
#pragma unroll
    for (int r = 0; r < 3; r++) J[3 * r + 2] += cm1() * d[r];
    if constexpr (PHYS != PHYS_COMPAT) {
      R v6 = e1[0] * (e2[1] * d[2] - e2[2] * d[1]) - e1[1] * (e2[0] * d[2] - e2[2] * d[0]) +
             e1[2] * (e2[0] * d[1] - e2[1] * d[0]);
      ea.mindt = fmin(ea.mindt, corner_dt<R>(e1, e2, d, v6, p.wave_speed, kTet10LengthFactor));
    }
    rel(4, d);
#pragma unroll
    for (int r = 0; r < 3; r++) {
      J[3 * r] += b * d[r];
      J[3 * r + 1] += a * d[r];
    }
    rel(5, d);
#pragma unroll
    for (int r = 0; r < 3; r++) {
      J[3 * r] -= b * d[r];
      J[3 * r + 1] += lb() * d[r];
      J[3 * r + 2] -= b * d[r];
    }
    rel(6, d);
#pragma unroll
    for (int r = 0; r < 3; r++) {
      J[3 * r] += la() * d[r];
      J[3 * r + 1] -= a * d[r];
      J[3 * r + 2] -= a * d[r];
    }
    rel(7, d);
#pragma unroll
    for (int r = 0; r < 3; r++) {
      J[3 * r] += c * d[r];
      J[3 * r + 2] += a * d[r];
    }
    rel(8, d);
#pragma unroll
    for (int r = 0; r < 3; r++) {
      J[3 * r + 1] += c * d[r];
      J[3 * r + 2] += b * d[r];
    }
    rel(9, d);
#pragma unroll
    for (int r = 0; r < 3; r++) {
      J[3 * r] -= c * d[r];
      J[3 * r + 1] -= c * d[r];
      J[3 * r + 2] += lc() * d[r];
    }
  }
  R Ji[9];
  const R detJ = inv3<R>(J, Ji);
  if (!(detJ > R(0))) ea.bad += 1.0;
  // Lv[r][k] = d v_r / d xi_k
  R Lv[9] = {R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0), R(0)};
  {
    R u[3];
    vel(0, u);
#pragma unroll
    for (int r = 0; r < 3; r++) Lv[3 * r] += am1() * u[r];
    vel(1, u);
#pragma unroll
    for (int r = 0; r < 3; r++) Lv[3 * r + 1] += bm1() * u[r];
    vel(2, u);
#pragma unroll
    for (int r = 0; r < 3; r++) {
      Lv[3 * r] += l1() * u[r];
      Lv[3 * r + 1] += l1() * u[r];
      Lv[3 * r + 2] += l1() * u[r];
    }
    vel(3, u);
#pragma unroll
    for (int r = 0; r < 3; r++) Lv[3 * r + 2] += cm1() * u[r];
    vel(4, u);
#pragma unroll
    for (int r = 0; r < 3; r++) {
      Lv[3 * r] += b * u[r];
      Lv[3 * r + 1] += a * u[r];
    }
    vel(5, u);
#pragma unroll
    for (int r = 0; r < 3; r++) {
      Lv[3 * r] -= b * u[r];
      Lv[3 * r + 1] += lb() * u[r];
      Lv[3 * r + 2] -= b * u[r];
    }
    vel(6, u);
#pragma unroll
    for (int r = 0; r < 3; r++) {
      Lv[3 * r] += la() * u[r];
      Lv[3 * r + 1] -= a * u[r];
      Lv[3 * r + 2] -= a * u[r];
    }
    vel(7, u);
#pragma unroll
    for (int r = 0; r < 3; r++) {
      Lv[3 * r] += c * u[r];
      Lv[3 * r + 2] += a * u[r];
    }
    vel(8, u);
#pragma unroll
    for (int r = 0; r < 3; r++) {
      Lv[3 * r + 1] += c * u[r];
      Lv[3 * r + 2] += b * u[r];
    }
    vel(9, u);
#pragma unroll
    for (int r = 0; r < 3; r++) {
      Lv[3 * r] -= c * u[r];
      Lv[3 * r + 1] -= c * u[r];
      Lv[3 * r + 2] += lc() * u[r];
    }
  }
  // grad v[r][s] = sum_k Lv[r][k] Ji[k][s]; engineering shear
  auto gv = [&](int r, int s) {
    return Lv[3 * r] * Ji[s] + Lv[3 * r + 1] * Ji[3 + s] + Lv[3 * r + 2] * Ji[6 + s];
  };
  R de[6] = {gv(0, 0), gv(1, 1), gv(2, 2), gv(0, 1) + gv(1, 0), gv(1, 2) + gv(2, 1),
             gv(0, 2) + gv(2, 0)};
#pragma unroll
  for (int k = 0; k < 6; k++) de[k] *= R(dt);
  const R wd = R(Q::w(q)) * detJ;
  R s[6];
  point_stress<R, PHYS, true, STREAM, DEFER>(p, e, q, de, wd, s, ea, &pre);
  // M[r][k] = sum_t (w detJ sigma)[r][t] Ji[k][t], sigma rows (s0 s3 s5),
  // (s3 s1 s4), (s5 s4 s2)
  R M[9], S[3];
  {
    R w[6];
#pragma unroll
    for (int k = 0; k < 6; k++) w[k] = wd * s[k];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      M[k] = w[0] * Ji[3 * k] + w[3] * Ji[3 * k + 1] + w[5] * Ji[3 * k + 2];
      M[3 + k] = w[3] * Ji[3 * k] + w[1] * Ji[3 * k + 1] + w[4] * Ji[3 * k + 2];
      M[6 + k] = w[5] * Ji[3 * k] + w[4] * Ji[3 * k + 1] + w[2] * Ji[3 * k + 2];
    }
#pragma unroll
    for (int r = 0; r < 3; r++) S[r] = M[3 * r] + M[3 * r + 1] + M[3 * r + 2];
  }
  // f_i = M G_i (f_5, f_6, f_9 through S = M (1, 1, 1)), summed over the
  // element's points
  auto emit = [&](int i, R f0, R f1, R f2) {
    const R f[3] = {quad_sum(f0), quad_sum(f1), quad_sum(f2)};
    put(i, f);
  };
  emit(0, am1() * M[0], am1() * M[3], am1() * M[6]);
  emit(1, bm1() * M[1], bm1() * M[4], bm1() * M[7]);
  emit(2, l1() * S[0], l1() * S[1], l1() * S[2]);
  emit(3, cm1() * M[2], cm1() * M[5], cm1() * M[8]);
  emit(4, b * M[0] + a * M[1], b * M[3] + a * M[4], b * M[6] + a * M[7]);
  emit(5, l * M[1] - b * S[0], l * M[4] - b * S[1], l * M[7] - b * S[2]);
  emit(6, l * M[0] - a * S[0], l * M[3] - a * S[1], l * M[6] - a * S[2]);
  emit(7, c * M[0] + a * M[2], c * M[3] + a * M[5], c * M[6] + a * M[8]);
  emit(8, c * M[1] + b * M[2], c * M[4] + b * M[5], c * M[7] + b * M[8]);
  emit(9, l * M[2] - c * S[0], l * M[5] - c * S[1], l * M[8] - c * S[2]);
}

// tet10 with the corner forces summed in registers, handed to put() at the end
template <int RULE, typename R, int PHYS, class Get, class Put>
__device__ __forceinline__ void tet10_core(const Params &p, const Get &get, double dt, int e,
                                           const Put &put, EAcc &ea) {
  R f[10][3];
#pragma unroll
  for (int i = 0; i < 10; i++) f[i][0] = f[i][1] = f[i][2] = R(0);
  auto add = [&](int i, const R dn[3], const R s[6], R wd) { add_bts<R>(f[i], dn, s, wd); };
  tet10_points<RULE, R, PHYS>(p, get, dt, e, add, ea);
#pragma unroll
  for (int i = 0; i < 10; i++) put(i, f[i]);
}

template <int NPE, int RULE, typename R, int PHYS, class Get, class Put>
__device__ __forceinline__ void elem_core(const Params &p, const Get &get, double dt, int e,
                                          const Put &put, EAcc &ea) {
  if constexpr (NPE == 4)
    tet4_core<R, PHYS, false>(p, get, dt, e, put, ea);
  else
    tet10_core<RULE, R, PHYS>(p, get, dt, e, put, ea);
}

// ---------------------------------------------------------------------------
// lumped mass (FP64): ROWSUM restates analysis.h:136-162 with the rule's
// points; HRZ gives rho V/4 (tet4), rho V/36 and 4 rho V/27 (tet10)
// ---------------------------------------------------------------------------
template <int NPE, int RULE>
__device__ __forceinline__ void elem_mass(const double (&xs)[NPE][3], int kind, double rho,
                                          double (&me)[NPE]) {
  double e1[3], e2[3], e3[3];
  const int o = NPE == 4 ? 0 : 2, i1 = NPE == 4 ? 1 : 0, i2 = NPE == 4 ? 2 : 1;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    e1[c] = xs[i1][c] - xs[o][c];
    e2[c] = xs[i2][c] - xs[o][c];
    e3[c] = xs[3][c] - xs[o][c];
  }
  double v6 = e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
              e1[2] * (e2[0] * e3[1] - e2[1] * e3[0]);
  if (kind == 1) {
    double M = rho * v6 / 6.0;
    if constexpr (NPE == 4) {
#pragma unroll
      for (int i = 0; i < 4; i++) me[i] = 0.25 * M;
    } else {
#pragma unroll
      for (int i = 0; i < 4; i++) me[i] = M / 36.0;
#pragma unroll
      for (int i = 4; i < 10; i++) me[i] = 4.0 * M / 27.0;
    }
    return;
  }
  // row sum; straight-edged elements have a constant Jacobian, det = v6
  using Q = Quad<RULE>;
#pragma unroll
  for (int i = 0; i < NPE; i++) me[i] = 0.0;
#pragma unroll
  for (int q = 0; q < Q::n; q++) {
    const double x = Q::pt(q, 0), y = Q::pt(q, 1), z = Q::pt(q, 2);
    double N[NPE];
    if constexpr (NPE == 4) {
      N[0] = 1.0 - x - y - z;
      N[1] = x;
      N[2] = y;
      N[3] = z;
    } else {
      double L1 = 1.0 - x - y - z, L2 = x, L3 = y, L4 = z;
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
#pragma unroll
    for (int i = 0; i < NPE; i++) me[i] += N[i] * Q::w(q) * v6 * rho;
  }
}

}  // namespace cpi
