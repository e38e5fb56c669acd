"""Pins the CPU oracle: closed forms, invariants, brute force, and the values
the reference's FEA_debug driver prints (tests/golden/fea_debug_values.json).
None of these retypes the oracle's formulas: each check is an independent
route to the same number."""
import json
import math
import os

import numpy as np
import pytest

import cppimpact_amd as ci
from oracle import pyoracle as O

E, NU, RHO = 68.9e9, 0.33, 2700.0
LAM = E * NU / ((1 + NU) * (1 - 2 * NU))
MU = E / (2 * (1 + NU))

# reference-coordinate positions of the tet10 nodes in the reference's
# numbering: corners 0..3 at xi=1, eta=1, origin, zeta=1; mid-sides of
# (0,1) (1,2) (0,2) (0,3) (1,3) (2,3)
CORNERS10 = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 0], [0, 0, 1]], float)
EDGES10 = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
NODES10 = np.vstack([CORNERS10] + [0.5 * (CORNERS10[a] + CORNERS10[b]) for a, b in EDGES10])
NODES4 = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)


def monomial_integral(a, b, c):
    return math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)


def lame_D():
    D = np.zeros((6, 6))
    D[:3, :3] = LAM
    D[[0, 1, 2], [0, 1, 2]] += 2 * MU
    D[[3, 4, 5], [3, 4, 5]] = MU
    return D


def tet10_from_corners(P):
    P = np.asarray(P, float)
    return np.vstack([P] + [0.5 * (P[a] + P[b]) for a, b in EDGES10])


def rand_tet(rng, npe):
    while True:
        P = rng.normal(size=(4, 3))
        if np.linalg.det(np.array([P[1] - P[0], P[2] - P[0], P[3] - P[0]])) > 0.3:
            break
    if npe == 4:
        return P
    # the quadratic corner map is an even permutation of the linear one
    return tet10_from_corners(P)


@pytest.mark.parametrize("rule,degree,tol", [
    (O.Q_EXACT4, 2, 1e-15), (O.Q_REF5, 3, 1e-9), (O.Q_REF4, 2, 1e-7), (O.Q_CENTROID, 1, 1e-15)])
def test_quadrature_integrates_monomials(rule, degree, tol):
    pts, w = O.quad_points(rule)
    for a in range(degree + 1):
        for b in range(degree + 1 - a):
            for c in range(degree + 1 - a - b):
                got = np.sum(w * pts[:, 0] ** a * pts[:, 1] ** b * pts[:, 2] ** c)
                assert abs(got - monomial_integral(a, b, c)) < tol * 10


@pytest.mark.parametrize("npe,nodes", [(4, NODES4), (10, NODES10)])
def test_basis_kronecker_partition_and_gradient(npe, nodes):
    for j, xj in enumerate(nodes):
        N = O.basis(npe, xj)
        assert np.allclose(N, np.eye(npe)[j], atol=1e-14)
    rng = np.random.default_rng(1)
    for _ in range(20):
        p = rng.dirichlet(np.ones(4))[:3]
        N = O.basis(npe, p)
        G = O.basis_grad(npe, p)
        assert abs(N.sum() - 1) < 1e-14
        assert np.abs(G.sum(axis=0)).max() < 1e-13
        h = 1e-6
        for q in range(3):
            dp = np.zeros(3)
            dp[q] = h
            fd = (O.basis(npe, p + dp) - O.basis(npe, p - dp)) / (2 * h)
            assert np.abs(fd - G[:, q]).max() < 1e-8


def test_inv3x3_against_numpy():
    rng = np.random.default_rng(2)
    for _ in range(10):
        A = rng.normal(size=(3, 3)) + 2 * np.eye(3)
        det, Ai = O.inv3x3(A)
        assert abs(det - np.linalg.det(A)) < 1e-12 * abs(det) * 10
        assert np.allclose(Ai, np.linalg.inv(A), rtol=1e-12, atol=1e-14)


def test_dmatrix_equals_lame_form():
    assert np.allclose(O.dmatrix(E, NU), lame_D(), rtol=1e-14)


def numpy_stiffness(npe, rule, xe):
    """K_e from numpy: J by finite differences of the mapped position, B
    from np.linalg.inv, D from the Lame constants."""
    pts, w = O.quad_points(rule)
    K = np.zeros((3 * npe, 3 * npe))
    D = lame_D()
    X = np.asarray(xe).reshape(npe, 3)
    for p, wq in zip(pts, w):
        G = np.zeros((npe, 3))
        h = 1e-5
        for q in range(3):
            dp = np.zeros(3)
            dp[q] = h
            G[:, q] = (O.basis(npe, p + dp) - O.basis(npe, p - dp)) / (2 * h)
        J = X.T @ G
        dN = G @ np.linalg.inv(J)
        B = np.zeros((6, 3 * npe))
        for i in range(npe):
            dx, dy, dz = dN[i]
            B[0, 3 * i] = dx
            B[1, 3 * i + 1] = dy
            B[2, 3 * i + 2] = dz
            B[3, 3 * i], B[3, 3 * i + 1] = dy, dx
            B[4, 3 * i + 1], B[4, 3 * i + 2] = dz, dy
            B[5, 3 * i], B[5, 3 * i + 2] = dz, dx
        K += wq * np.linalg.det(J) * B.T @ D @ B
    return K


@pytest.mark.parametrize("npe,rule", [(4, O.Q_REF5), (10, O.Q_REF5), (10, O.Q_EXACT4)])
def test_stiffness_against_independent_numpy_build(npe, rule):
    rng = np.random.default_rng(3)
    xe = rand_tet(rng, npe)
    K = O.element_stiffness(npe, rule, E, NU, xe.reshape(-1))
    Kn = numpy_stiffness(npe, rule, xe)
    assert np.abs(K - Kn).max() < 1e-7 * np.abs(K).max()
    du = rng.normal(size=3 * npe)
    f = O.element_force_compat(npe, rule, E, NU, xe.reshape(-1), du)
    assert np.allclose(f, Kn @ du, rtol=1e-6, atol=1e-7 * np.abs(Kn @ du).max())


@pytest.mark.parametrize("npe", [4, 10])
def test_stiffness_symmetric_psd_with_rigid_null_space(npe):
    rng = np.random.default_rng(4)
    xe = rand_tet(rng, npe)
    K = O.element_stiffness(npe, O.Q_REF5 if npe == 4 else O.Q_EXACT4, E, NU, xe.reshape(-1))
    assert np.abs(K - K.T).max() < 1e-12 * np.abs(K).max()
    ev = np.linalg.eigvalsh(0.5 * (K + K.T))
    scale = ev.max()
    assert np.sum(np.abs(ev) < 1e-10 * scale) == 6
    assert ev.min() > -1e-10 * scale
    X = xe.reshape(npe, 3)
    for mode in range(6):
        u = np.zeros_like(X)
        if mode < 3:
            u[:, mode] = 1.0
        else:
            axis = np.eye(3)[mode - 3]
            u = np.cross(axis, X)
        assert np.abs(K @ u.reshape(-1)).max() < 1e-9 * scale * np.abs(u).max()


def test_patch_test_linear_field_gives_zero_interior_force():
    x, conn = ci.box_mesh(3, 3, 3, 1.0, 1.0, 1.0, jitter=0.2, seed=5)
    Gr = np.array([[1e-3, 2e-4, -3e-4], [5e-4, -1e-3, 1e-4], [2e-4, 3e-4, 7e-4]])
    u = x @ Gr.T
    f = np.zeros_like(x)
    for c in conn:
        fe = O.element_force_compat(4, O.Q_REF5, E, NU, x[c].reshape(-1), u[c].reshape(-1))
        f[c] += fe.reshape(4, 3)
    interior = np.all((x > 1e-9) & (x < 1 - 1e-9), axis=1)
    assert interior.any()
    assert np.abs(f[interior]).max() < 1e-9 * np.abs(f).max()


def closed_form_energy(slope, volume):
    """1/2 V eps.D.eps for u = slope * x * (1, -nu, -nu): the field depends on
    x only, so eps = slope * (1, 0, 0, -nu, 0, -nu) (Voigt, engineering
    shear)."""
    e = slope * np.array([1.0, 0, 0, -NU, 0, -NU])
    return 0.5 * volume * e @ lame_D() @ e


def test_fea_debug_known_answers(golden):
    cases = json.load(open(os.path.join(golden, "fea_debug_values.json")))["cases"]
    a = 0.01
    for c in cases:
        npe = 4 if c["basis"] == "linear" else 10
        mesh = ci.Mesh(npe)
        assert mesh.load_mesh(os.path.join(golden, c["mesh"])) == 0
        W, V = O.debug_strain(npe, O.Q_REF5, E, NU, mesh.xloc, mesh.element_nodes, a, c["def"])
        # the reference's printed values
        assert abs(W - c["W"]) <= 5e-15 * abs(c["W"]) + 1e-6, (c, W)
        assert abs(V - c["V"]) <= 1e-6
        # closed forms (5-point weights are truncated at 1e-11)
        if c["def"] == 0:
            ref = closed_form_energy(a, V)
            assert abs(W - ref) < 1e-9 * ref
        elif c["mesh"] == "linear_tet.inp":
            # the nodal interpolant of x^2/2 over [0, 10] has slope 5 a
            ref = closed_form_energy(5 * a, V)
            assert abs(W - ref) < 1e-9 * ref
        elif c["mesh"] == "quad_tet.inp":
            # tet10 reproduces the quadratic field: eps = a x (1,0,0,-nu,0,-nu)
            ix2 = 10.0 ** 5 * monomial_integral(2, 0, 0)
            ref = closed_form_energy(a, ix2)
            assert abs(W - ref) < 1e-9 * ref
        else:
            # cube mesh, quadratic field: converges to the exact energy
            lo, hi = mesh.xloc[:, 0].min(), mesh.xloc[:, 0].max()
            area = V / (hi - lo)
            ix2 = area * (hi ** 3 - lo ** 3) / 3
            ref = closed_form_energy(a, ix2)
            assert abs(W - ref) < 2e-3 * ref


def test_lumped_masses_closed_forms():
    rng = np.random.default_rng(6)
    P = rand_tet(rng, 4)
    V = np.linalg.det(np.array([P[1] - P[0], P[2] - P[0], P[3] - P[0]])) / 6
    m = O.element_mass(4, O.Q_REF5, O.MASS_ROWSUM, RHO, P.reshape(-1))
    assert np.allclose(m, RHO * V / 4, rtol=1e-9)
    m = O.element_mass(4, O.Q_CENTROID, O.MASS_ROWSUM, RHO, P.reshape(-1))
    assert np.allclose(m, RHO * V / 4, rtol=1e-14)
    X10 = tet10_from_corners(P)
    m = O.element_mass(10, O.Q_REF5, O.MASS_ROWSUM, RHO, X10.reshape(-1))
    assert np.allclose(m[:4], -RHO * V / 20, rtol=1e-8)
    assert np.allclose(m[4:], RHO * V / 5, rtol=1e-8)
    m = O.element_mass(10, O.Q_EXACT4, O.MASS_HRZ, RHO, X10.reshape(-1))
    assert np.allclose(m[:4], RHO * V / 36, rtol=1e-13)
    assert np.allclose(m[4:], 4 * RHO * V / 27, rtol=1e-13)
    assert abs(m.sum() - RHO * V) < 1e-13 * RHO * V


def test_radial_return_pure_shear_closed_form():
    sy, H = 276e6, 700e6
    tau_y = sy / math.sqrt(3)
    gamma = 3 * tau_y / MU  # three times the yield strain
    s, ep, dg = O.radial_return(E, NU, sy, H, [0, 0, 0, gamma, 0, 0], np.zeros(6), 0.0)
    # tau = sy/sqrt3 + H (gamma - tau/G)/3  (gamma_p = sqrt3 ep)
    tau = (tau_y + H * gamma / 3) / (1 + H / (3 * MU))
    assert abs(s[3] - tau) < 1e-9 * tau
    assert abs(ep - (gamma - tau / MU) / math.sqrt(3)) < 1e-9 * ep
    assert np.abs(s[[0, 1, 2, 4, 5]]).max() < 1e-6
    # elastic below yield, and a pure volumetric strain never yields
    s, ep, dg = O.radial_return(E, NU, sy, H, [0, 0, 0, 0.5 * tau_y / MU, 0, 0], np.zeros(6), 0.0)
    assert dg == 0 and ep == 0 and abs(s[3] - 0.5 * tau_y) < 1e-6
    s, ep, dg = O.radial_return(E, NU, sy, H, [0.01, 0.01, 0.01, 0, 0, 0], np.zeros(6), 0.0)
    assert dg == 0 and abs(s[0] - (3 * LAM + 2 * MU) * 0.01) < 1e-3


def test_radial_return_consistency_on_random_paths():
    rng = np.random.default_rng(7)
    sy, H = 276e6, 700e6
    for _ in range(50):
        s0 = rng.normal(scale=150e6, size=6)
        ep0 = abs(rng.normal(scale=1e-3))
        de = rng.normal(scale=3e-3, size=6)
        s, ep, dg = O.radial_return(E, NU, sy, H, de, s0, ep0)
        d = s.copy()
        d[:3] -= d[:3].mean()
        nrm = math.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2 + 2 * (d[3] ** 2 + d[4] ** 2 + d[5] ** 2))
        lim = math.sqrt(2 / 3) * (sy + H * ep)
        if dg > 0:
            assert abs(nrm - lim) < 1e-8 * lim
            assert abs(ep - ep0 - math.sqrt(2 / 3) * dg) < 1e-15
        else:
            assert nrm <= lim * (1 + 1e-12)
        # the pressure is elastic either way
        assert abs(s[:3].mean() - s0[:3].mean() - (LAM + 2 * MU / 3) * de[:3].sum()) < 1e-4 * (
            abs(s[:3].mean()) + 1)


def python_greedy(conn, nn):
    used = [set() for _ in range(nn)]
    col = []
    for c in conn:
        taken = set().union(*(used[n] for n in c))
        k = 0
        while k in taken:
            k += 1
        col.append(k)
        for n in c:
            used[n].add(k)
    return np.array(col)


@pytest.mark.parametrize("quadratic,jitter", [(False, 0.0), (False, 0.2), (True, 0.0)])
def test_colouring_valid_greedy_and_brute_force(quadratic, jitter):
    x, conn = ci.box_mesh(4, 3, 3, 1.0, 1.0, 1.0, quadratic=quadratic, jitter=jitter, seed=9)
    col, n = O.greedy_coloring(conn, x.shape[0])
    assert np.array_equal(col, python_greedy(conn, x.shape[0]))
    assert n == col.max() + 1
    for c in range(n):
        nodes = conn[col == c].reshape(-1)
        assert len(np.unique(nodes)) == nodes.size  # no shared node inside a colour


def test_free_flight_is_exact_leapfrog():
    x, conn = ci.box_mesh(2, 2, 2, 0.1, 0.1, 0.1)
    # dt well inside the stable range (element limit ~5.8e-6 s here)
    dt, n, v0 = 1e-6, 30, np.array([1.0, -2.0, 3.0])
    for phys in (O.COMPAT, O.EP):
        out = O.Solver(x, conn, v0, E=E, nu=NU, rho=RHO, physics=phys, dt=dt).solve(n)
        t = n * dt
        g = np.array([0, 0, -9.81])
        assert np.abs(out["x"] - (x + v0 * t + 0.5 * g * t * t)).max() < 1e-12
        assert np.abs(out["v_i"] - (v0 + g * t)).max() < 1e-9


def test_single_node_wall_penetration_acceleration():
    X = np.array([[0, 0, -1e-4], [1, 0, 0.1], [0, 1, 0.1], [0, 0, 1.1]], float)
    conn = np.array([[0, 1, 2, 3]])
    k = 1e9
    s = O.Solver(X, conn, (0, 0, 0), E=E, nu=NU, rho=RHO, physics=O.COMPAT,
                 wall=dict(location=0.0, stiffness=k), dt=1e-9)
    out = s.solve(0)
    m0 = out["m"][0]
    V = np.linalg.det(np.array([X[1] - X[0], X[2] - X[0], X[3] - X[0]])) / 6
    assert abs(m0 - RHO * V / 4) < 1e-9 * m0
    # a = -d k / m - g on the penetrating node, gravity elsewhere
    assert abs(out["a"][0, 2] - (1e-4 * k / m0 - 9.81)) < 1e-9 * 1e-4 * k / m0
    assert np.allclose(out["a"][1:, 2], -9.81)


def test_two_steps_by_hand_on_one_element():
    rng = np.random.default_rng(8)
    X = rand_tet(rng, 4) * 0.01
    conn = np.array([[0, 1, 2, 3]])
    v0 = rng.normal(size=(4, 3))
    dt = 1e-7
    out = O.Solver(X, conn, v0, E=E, nu=NU, rho=RHO, physics=O.COMPAT, dt=dt).solve(2)
    K = numpy_stiffness(4, O.Q_REF5, X)
    m = np.full(4, RHO * np.linalg.det(np.array([X[1] - X[0], X[2] - X[0], X[3] - X[0]])) / 24)
    g = np.array([0, 0, -9.81])
    x, v = X.copy(), v0 + 0.5 * dt * g  # a0 = g (du = 0)
    for _ in range(2):
        du = dt * v
        a = -(K @ du.reshape(-1)).reshape(4, 3) / m[:, None] + g
        x, v = x + du, v + dt * a
    assert np.abs(out["x"] - x).max() < 1e-12 * np.abs(x).max()
    assert np.abs(out["v"] - v).max() < 1e-6 * np.abs(v).max()


def test_cfl_kuhn_closed_form_and_element_frequency_bound():
    c = math.sqrt((LAM + 2 * MU) / RHO)
    a = 0.025
    x, conn = ci.box_mesh(1, 1, 1, a, a, a)
    # Kuhn tetrahedron: minimum altitude a/sqrt(2), tet4 length factor 2
    for e in conn:
        assert abs(O.element_dt(4, E, NU, RHO, x[e].reshape(-1)) - a / math.sqrt(2) / c / 2) < 1e-18

    def limit(npe, rule, xe):
        K = O.element_stiffness(npe, rule, E, NU, xe)
        m = O.element_mass(npe, rule, O.MASS_HRZ, RHO, xe)
        mi = np.repeat(1 / np.sqrt(m), 3)
        return 2 / math.sqrt(np.linalg.eigvalsh(mi[:, None] * K * mi[None, :]).max())

    rng = np.random.default_rng(10)
    for npe, rule in ((4, O.Q_CENTROID), (10, O.Q_EXACT4)):
        xs, cs = ci.box_mesh(2, 2, 2, a, a, a, quadratic=npe == 10, jitter=0.25, seed=11)
        for e in cs[rng.choice(len(cs), 12, replace=False)]:
            xe = xs[e].reshape(-1)
            assert O.element_dt(npe, E, NU, RHO, xe) <= limit(npe, rule, xe)
        # arbitrary shapes, including slivers
        for _ in range(300):
            P = rng.normal(size=(4, 3))
            d = np.linalg.det(np.array([P[1] - P[0], P[2] - P[0], P[3] - P[0]]))
            if abs(d) < 1e-3:
                continue
            if d < 0:
                P[[1, 2]] = P[[2, 1]]
            xe = (P if npe == 4 else tet10_from_corners(P)).reshape(-1)
            assert O.element_dt(npe, E, NU, RHO, xe) <= limit(npe, rule, xe)


def _balance(dt_frac, physics_time, with_wall):
    x, conn = ci.box_mesh(4, 4, 4, 0.25, 0.25, 0.25, origin=(0, 0, 1e-6), jitter=0.1, seed=2)
    # a vibrating block: rigid drop plus a shear velocity field
    v0 = np.zeros_like(x)
    v0[:, 2] = -5.0
    v0[:, 0] = 20.0 * (x[:, 2] - 0.125)
    wall = dict(location=0.0, stiffness=10 * E * 0.0625) if with_wall else None
    s = O.Solver(x, conn, v0, E=E, nu=NU, rho=RHO, physics=O.EP, wall=wall, dt=1.0)
    de, dc = s.stable_dt()
    dts = 1 / math.sqrt(1 / de ** 2 + (1 / dc ** 2 if with_wall else 0.0))
    s.p.dt = dt_frac * dts
    n = int(round(physics_time / s.p.dt))
    out = s.solve(n)
    en = out["energy"]
    assert out["epbar"].max() == 0.0  # stays elastic
    return np.abs(en[:, O.E_BALANCE]).max() / en[0, O.E_KINETIC]


def test_ep_energy_balance_converges_second_order():
    # leapfrog with the midpoint internal-work rule: the balance error of an
    # elastic run is O(dt^2); halving dt divides it by ~4
    T = 2e-5
    e1 = _balance(0.4, T, False)
    e2 = _balance(0.2, T, False)
    assert e1 < 1e-3
    assert 3.0 < e1 / e2 < 5.0
    # with the penalty wall engaged the error stays small at half the step
    assert _balance(0.5, T, True) < 2e-2


def test_stress_strain_of_a_linear_field_is_its_closed_form():
    """u = A x gives strain (A11, A22, A33, A12+A21, A23+A32, A13+A31) at any
    point and stress = D strain (engineering shear, Voigt xx,yy,zz,xy,yz,xz)."""
    rng = np.random.default_rng(7)
    A = rng.normal(scale=1e-3, size=(3, 3))
    want = np.array([A[0, 0], A[1, 1], A[2, 2], A[0, 1] + A[1, 0], A[1, 2] + A[2, 1],
                     A[0, 2] + A[2, 0]])
    D = np.zeros((6, 6))
    D[:3, :3] = LAM
    D[:3, :3] += 2 * MU * np.eye(3)
    D[3:, 3:] = MU * np.eye(3)
    for npe, nodes in ((4, NODES4), (10, NODES10)):
        xe = nodes @ np.array([[1.0, 0.2, 0.1], [0.0, 0.9, 0.3], [0.1, 0.0, 1.1]]).T
        ue = xe @ A.T
        for pt in ([0.25, 0.25, 0.25], [0.1, 0.2, 0.3]):
            st, ss = O.stress_strain(npe, E, NU, xe, ue, pt)
            assert np.allclose(st, want, rtol=1e-10, atol=1e-16)
            assert np.allclose(ss, D @ want, rtol=1e-9, atol=1e-3)


def test_ep_element_force_in_the_elastic_range_equals_ke_du():
    """From a stress-free state and an increment below yield, the EP force
    sum_q w detJ B^T (sigma_n + D B du) equals K_e du on the same rule."""
    rng = np.random.default_rng(8)
    for npe, nodes, rule in ((4, NODES4, O.Q_EXACT4), (4, NODES4, O.Q_CENTROID),
                             (10, NODES10, O.Q_EXACT4)):
        xe = nodes * 0.01 + rng.uniform(-5e-4, 5e-4, nodes.shape)
        du = rng.normal(scale=1e-9, size=3 * npe)
        nq = 1 if rule == O.Q_CENTROID else 4
        f, s, ep, _ = O.element_force_ep(npe, rule, E, NU, 276e6, 700e6, xe, du,
                                         np.zeros(6 * nq), np.zeros(nq))
        fk = O.element_force_compat(npe, rule, E, NU, xe, du)
        assert np.abs(ep).max() == 0.0
        assert np.allclose(f, fk, rtol=1e-9, atol=1e-12 * np.abs(fk).max())
