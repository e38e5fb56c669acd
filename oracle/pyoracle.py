"""ctypes front end of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point and by
bench.py's cpu_baseline leg, where it is the checker or the timed CPU
baseline.  The product package never imports it.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

Q_REF5, Q_REF4, Q_EXACT4, Q_CENTROID = 0, 1, 2, 3
MASS_ROWSUM, MASS_HRZ = 0, 1
COMPAT, EP = 0, 1
E_KINETIC, E_INTERNAL, E_PLASTIC, E_PENALTY, E_GRAVITY, E_CONTACTF, E_BALANCE, E_MINDT = range(8)
NENERGY = 8

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class Problem(C.Structure):
    _fields_ = [
        ("num_nodes", C.c_int), ("num_elements", C.c_int), ("npe", C.c_int),
        ("X0", _dp), ("conn", _ip), ("v0", _dp),
        ("E", C.c_double), ("nu", C.c_double), ("rho", C.c_double),
        ("sigma_y", C.c_double), ("H", C.c_double),
        ("physics", C.c_int), ("quad_rule", C.c_int), ("mass_kind", C.c_int),
        ("wall_enabled", C.c_int), ("wall_dim", C.c_int), ("wall_normal", C.c_int),
        ("wall_location", C.c_double), ("wall_stiffness", C.c_double),
        ("gravity", C.c_double * 3), ("dt", C.c_double),
        ("contact_mask", C.POINTER(C.c_ubyte)), ("fixed_mask", C.POINTER(C.c_ubyte)),
    ]


class State(C.Structure):
    _fields_ = [(n, _dp) for n in ("x", "v", "a", "m", "v_i", "strain", "stress",
                                   "sig", "epbar", "energy", "Xref")]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_quad_npts.restype = C.c_int
        L.orc_quad_point.restype = C.c_double
        L.orc_quad_point.argtypes = [C.c_int, C.c_int, _dp]
        L.orc_basis.argtypes = [C.c_int, _dp, _dp]
        L.orc_basis_grad.argtypes = [C.c_int, _dp, _dp]
        L.orc_inv3x3.restype = C.c_double
        L.orc_inv3x3.argtypes = [_dp, _dp]
        L.orc_dmatrix.argtypes = [C.c_double, C.c_double, _dp]
        L.orc_element_stiffness.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.orc_element_force_compat.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp]
        L.orc_element_mass.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp]
        L.orc_element_volume.restype = C.c_double
        L.orc_element_volume.argtypes = [C.c_int, C.c_int, _dp]
        L.orc_strain_energy.restype = C.c_double
        L.orc_strain_energy.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp]
        L.orc_stress_strain.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp]
        L.orc_element_dt.restype = C.c_double
        L.orc_element_dt.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _dp]
        L.orc_radial_return.restype = C.c_double
        L.orc_radial_return.argtypes = [C.c_double] * 4 + [_dp, _dp, _dp]
        L.orc_element_force_ep.argtypes = [C.c_int, C.c_int] + [C.c_double] * 4 + [_dp] * 6
        L.orc_debug_strain.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                       C.c_int, _ip, _dp, C.c_double, C.c_int, _dp]
        L.orc_greedy_coloring.restype = C.c_int
        L.orc_greedy_coloring.argtypes = [C.c_int, C.c_int, _ip, C.c_int, _ip]
        L.orc_solve.restype = C.c_int
        L.orc_solve.argtypes = [C.POINTER(Problem), C.c_int, C.POINTER(State)]
        L.orc_stable_dt.restype = C.c_double
        L.orc_stable_dt.argtypes = [C.POINTER(Problem), _dp]
        L.orc_solve_timed.restype = C.c_int
        L.orc_solve_timed.argtypes = [C.POINTER(Problem), C.c_int, C.POINTER(State), _dp]
        L.orc_solve_threaded.restype = C.c_int
        L.orc_solve_threaded.argtypes = [C.POINTER(Problem), C.c_int, C.POINTER(State), C.c_int,
                                         _dp]
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def quad_points(rule):
    n = lib().orc_quad_npts(rule)
    pts, ws = np.zeros((n, 3)), np.zeros(n)
    for k in range(n):
        p = np.zeros(3)
        ws[k] = lib().orc_quad_point(rule, k, _d(p))
        pts[k] = p
    return pts, ws


def basis(npe, pt):
    N = np.zeros(npe)
    lib().orc_basis(npe, _d(f64(pt)), _d(N))
    return N


def basis_grad(npe, pt):
    G = np.zeros(3 * npe)
    lib().orc_basis_grad(npe, _d(f64(pt)), _d(G))
    return G.reshape(npe, 3)


def inv3x3(A):
    A = f64(A).reshape(9)
    Ai = np.zeros(9)
    det = lib().orc_inv3x3(_d(A), _d(Ai))
    return det, Ai.reshape(3, 3)


def dmatrix(E, nu):
    D = np.zeros(36)
    lib().orc_dmatrix(E, nu, _d(D))
    return D.reshape(6, 6)


def element_stiffness(npe, rule, E, nu, xe):
    K = np.zeros((3 * npe, 3 * npe))
    lib().orc_element_stiffness(npe, rule, E, nu, _d(f64(xe)), _d(K))
    return K


def element_force_compat(npe, rule, E, nu, xe, due):
    f = np.zeros(3 * npe)
    lib().orc_element_force_compat(npe, rule, E, nu, _d(f64(xe)), _d(f64(due)), _d(f))
    return f


def element_mass(npe, rule, kind, rho, xe):
    m = np.zeros(npe)
    lib().orc_element_mass(npe, rule, kind, rho, _d(f64(xe)), _d(m))
    return m


def element_volume(npe, rule, xe):
    return lib().orc_element_volume(npe, rule, _d(f64(xe)))


def strain_energy(npe, rule, E, nu, xe, ue):
    return lib().orc_strain_energy(npe, rule, E, nu, _d(f64(xe)), _d(f64(ue)))


def stress_strain(npe, E, nu, xe, ue, pt):
    st, ss = np.zeros(6), np.zeros(6)
    lib().orc_stress_strain(npe, E, nu, _d(f64(xe)), _d(f64(ue)), _d(f64(pt)), _d(st), _d(ss))
    return st, ss


def element_dt(npe, E, nu, rho, xe):
    return lib().orc_element_dt(npe, E, nu, rho, _d(f64(xe)))


def radial_return(E, nu, sy0, H, deps, sig, epbar):
    s = f64(sig).copy()
    ep = np.array([float(epbar)])
    dg = lib().orc_radial_return(E, nu, sy0, H, _d(f64(deps)), _d(s), _d(ep))
    return s, ep[0], dg


def element_force_ep(npe, rule, E, nu, sy0, H, xe, due, sig, epbar):
    nq = lib().orc_quad_npts(rule)
    s = f64(sig).reshape(nq, 6).copy()
    ep = f64(epbar).reshape(nq).copy()
    f = np.zeros(3 * npe)
    w = np.zeros(2)
    lib().orc_element_force_ep(npe, rule, E, nu, sy0, H, _d(f64(xe)), _d(f64(due)),
                               _d(s), _d(ep), _d(f), _d(w))
    return f, s, ep, w


def debug_strain(npe, rule, E, nu, xloc, conn, alpha, case):
    xloc = f64(xloc).reshape(-1, 3)
    conn = np.ascontiguousarray(conn, dtype=np.int32).reshape(-1, npe)
    out = np.zeros(2)
    lib().orc_debug_strain(npe, rule, E, nu, xloc.shape[0], conn.shape[0], _i(conn),
                           _d(xloc), alpha, case, _d(out))
    return out[0], out[1]


def greedy_coloring(conn, num_nodes):
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    ne, npe = conn.shape
    col = np.zeros(ne, dtype=np.int32)
    n = lib().orc_greedy_coloring(ne, npe, _i(conn), num_nodes, _i(col))
    if n < 0:
        raise MemoryError("colouring")
    return col, n


class Solver:
    """Holds one problem and its state arrays for orc_solve."""

    def __init__(self, X0, conn, v0, *, E, nu, rho, sigma_y=276e6, H=700e6,
                 physics=COMPAT, quad_rule=None, mass_kind=None, wall=None,
                 gravity=(0.0, 0.0, -9.81), dt=1e-7, contact_nodes=None, fixed_nodes=None):
        """contact_nodes: node ids the wall acts on (None: all nodes);
        fixed_nodes: node ids held in place (zero initial velocity and
        acceleration, no contact)."""
        self.X0 = f64(X0).reshape(-1, 3)
        self.conn = np.ascontiguousarray(conn, dtype=np.int32)
        self.npe = self.conn.shape[1]
        self.v0 = f64(np.broadcast_to(f64(v0), self.X0.shape)).copy()
        if quad_rule is None:
            quad_rule = Q_REF5 if physics == COMPAT else (Q_CENTROID if self.npe == 4 else Q_EXACT4)
        if mass_kind is None:
            mass_kind = MASS_ROWSUM if physics == COMPAT else MASS_HRZ
        p = Problem()
        p.num_nodes, p.num_elements, p.npe = self.X0.shape[0], self.conn.shape[0], self.npe
        p.X0, p.conn, p.v0 = _d(self.X0), _i(self.conn), _d(self.v0)
        p.E, p.nu, p.rho, p.sigma_y, p.H = E, nu, rho, sigma_y, H
        p.physics, p.quad_rule, p.mass_kind = physics, quad_rule, mass_kind
        if wall is None:
            p.wall_enabled = 0
            p.wall_dim, p.wall_normal, p.wall_location, p.wall_stiffness = 2, 1, 0.0, 0.0
        else:
            p.wall_enabled = 1
            p.wall_dim = int(wall.get("dim", 2))
            p.wall_normal = int(wall.get("normal", 1))
            p.wall_location = float(wall["location"])
            p.wall_stiffness = float(wall["stiffness"])
        p.gravity = (C.c_double * 3)(*gravity)
        p.dt = dt
        self.masks = []
        for name, ids in (("contact_mask", contact_nodes), ("fixed_mask", fixed_nodes)):
            if ids is None:
                setattr(p, name, None)
                continue
            m = np.zeros(self.X0.shape[0], dtype=np.uint8)
            m[np.asarray(ids, dtype=np.int64)] = 1
            self.masks.append(m)
            setattr(p, name, m.ctypes.data_as(C.POINTER(C.c_ubyte)))
        self.p = p
        self.nq = lib().orc_quad_npts(quad_rule)

    def stable_dt(self):
        dc = C.c_double(0.0)
        de = lib().orc_stable_dt(C.byref(self.p), C.byref(dc))
        return de, dc.value

    def solve(self, nsteps, with_strain=False, with_energy=True, timed=False, threads=1):
        """threads > 1 runs orc_solve_threaded (bit-identical results)."""
        n, e = self.p.num_nodes, self.p.num_elements
        out = {k: np.zeros((n, 3)) for k in ("x", "v", "a", "v_i")}
        out["m"] = np.zeros(n)
        out["Xref"] = np.zeros((n, 3))
        out["sig"] = np.zeros((e, self.nq, 6))
        out["epbar"] = np.zeros((e, self.nq))
        if with_strain:
            out["strain"] = np.zeros((n, 6))
            out["stress"] = np.zeros((n, 6))
        if with_energy:
            out["energy"] = np.zeros((nsteps + 1, NENERGY))
        s = State()
        for k in ("x", "v", "a", "m", "v_i", "sig", "epbar", "Xref"):
            setattr(s, k, _d(out[k]))
        for k in ("strain", "stress", "energy"):
            setattr(s, k, _d(out[k]) if k in out else None)
        secs = C.c_double(0.0)
        if threads > 1:
            rc = lib().orc_solve_threaded(C.byref(self.p), nsteps, C.byref(s), int(threads),
                                          C.byref(secs))
        else:
            rc = lib().orc_solve_timed(C.byref(self.p), nsteps, C.byref(s), C.byref(secs))
        if rc != 0:
            raise RuntimeError("orc_solve failed")
        out["loop_seconds"] = secs.value
        return out
