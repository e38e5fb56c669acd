"""Python mirror of the reference's host classes, driving the HIP step.

The reference's public surface (Mesh, BaseMaterial/Elastoplastic, Wall,
Dynamics; include/mesh.h, basematerial.h, elastoplastic.h, wall.h,
dynamics.h) with the same names, argument meaning and state arrays.  Every
compute call goes through the C ABI of libcppimpact_hip.so; nothing here
computes a step on the CPU.
"""
import ctypes as C

import numpy as np

from . import _lib as L


class Mesh:
    """Mesh<T, nodes_per_element> (mesh.h:11-330)."""

    def __init__(self, nodes_per_element=4):
        if nodes_per_element not in (4, 10):
            raise ValueError("nodes_per_element must be 4 or 10")
        self.nodes_per_element = nodes_per_element
        self.num_elements = self.num_nodes = self.num_node_sets = 0
        self.num_fixed_nodes = self.num_slave_nodes = 0
        self.element_nodes = np.zeros((0, nodes_per_element), dtype=np.int32)
        self.xloc = np.zeros((0, 3))
        self.node_set_starts = np.zeros(0, dtype=np.int32)
        self.node_set_indices = np.zeros(0, dtype=np.int32)
        self.fixed_nodes = np.zeros(0, dtype=np.int32)
        self.slave_nodes = np.zeros(0, dtype=np.int32)
        self.node_set_names = []

    def load_mesh(self, filename):
        """mesh.h:57-267: returns 0 on success, 1 when the file cannot be
        opened; a malformed file raises L.CpiError."""
        cnt = L.InpCounts()
        rc = L.lib().cpi_inp_read(filename.encode(), self.nodes_per_element, C.byref(cnt),
                                  None, None, None, None, None, None, None)
        if rc == -6 and b"Failed to open" in (L.lib().cpi_last_error(None) or b""):
            return 1
        L.check(rc)
        npe = self.nodes_per_element
        x = np.zeros((cnt.num_nodes, 3))
        conn = np.zeros((cnt.num_elements, npe), dtype=np.int32)
        starts = np.zeros(cnt.num_node_sets + 1, dtype=np.int32)
        idx = np.zeros(max(cnt.num_set_indices, 1), dtype=np.int32)
        fixed = np.zeros(max(cnt.num_fixed_nodes, 1), dtype=np.int32)
        slave = np.zeros(max(cnt.num_slave_nodes, 1), dtype=np.int32)
        names = C.create_string_buffer(max(cnt.names_bytes, 1))
        L.check(L.lib().cpi_inp_read(filename.encode(), npe, C.byref(cnt), L.dptr(x),
                                     L.iptr(conn), L.iptr(starts), L.iptr(idx), L.iptr(fixed),
                                     L.iptr(slave), names))
        self.num_nodes, self.num_elements = cnt.num_nodes, cnt.num_elements
        self.num_node_sets = cnt.num_node_sets
        self.num_fixed_nodes, self.num_slave_nodes = cnt.num_fixed_nodes, cnt.num_slave_nodes
        self.xloc, self.element_nodes = x, conn
        self.node_set_starts = starts
        self.node_set_indices = idx[:cnt.num_set_indices]
        self.fixed_nodes = fixed[:cnt.num_fixed_nodes]
        self.slave_nodes = slave[:cnt.num_slave_nodes]
        raw = names.raw[:cnt.names_bytes]
        self.node_set_names = [s.decode() for s in raw.split(b"\0")[:cnt.num_node_sets]]
        return 0

    @classmethod
    def from_arrays(cls, xloc, element_nodes, slave_nodes=None, fixed_nodes=None):
        conn = np.ascontiguousarray(element_nodes, dtype=np.int32)
        m = cls(conn.shape[1])
        m.xloc = np.ascontiguousarray(xloc, dtype=np.float64).reshape(-1, 3)
        m.element_nodes = conn
        m.num_nodes, m.num_elements = m.xloc.shape[0], conn.shape[0]
        if slave_nodes is not None:
            m.slave_nodes = np.ascontiguousarray(slave_nodes, dtype=np.int32)
            m.num_slave_nodes = m.slave_nodes.size
        if fixed_nodes is not None:
            m.fixed_nodes = np.ascontiguousarray(fixed_nodes, dtype=np.int32)
            m.num_fixed_nodes = m.fixed_nodes.size
        return m

    @classmethod
    def box(cls, nx, ny, nz, lx, ly, lz, origin=(0.0, 0.0, 0.0), quadratic=False, jitter=0.0,
            seed=0):
        x, conn = L.box_mesh(nx, ny, nz, lx, ly, lz, origin, quadratic, jitter, seed)
        return cls.from_arrays(x, conn)

    @classmethod
    def cylinder(cls, ntheta, nz, nr, radius, length, thickness, origin=(0.0, 0.0, 0.0),
                 quadratic=False):
        x, conn = L.cylinder_mesh(ntheta, nz, nr, radius, length, thickness, origin, quadratic)
        return cls.from_arrays(x, conn)

    def write_inp(self, filename, slave_nodes=None):
        sl = self.slave_nodes if slave_nodes is None else np.asarray(slave_nodes, np.int32)
        d = self._desc(sl)
        L.check(L.lib().cpi_inp_write(filename.encode(), C.byref(d)))

    def _desc(self, slave=None):
        fixed = self.fixed_nodes if len(self.fixed_nodes) else None
        self._keep = (np.ascontiguousarray(self.xloc, dtype=np.float64),
                      np.ascontiguousarray(self.element_nodes, dtype=np.int32),
                      None if slave is None else np.ascontiguousarray(slave, dtype=np.int32),
                      None if fixed is None else np.ascontiguousarray(fixed, dtype=np.int32))
        x, conn, sl, fx = self._keep
        return L.MeshDesc(x.shape[0], conn.shape[0], conn.shape[1], L.dptr(x), L.iptr(conn),
                          0 if sl is None else sl.size, L.iptr(sl) if sl is not None else None,
                          0 if fx is None else fx.size, L.iptr(fx) if fx is not None else None)


class BaseMaterial:
    """BaseMaterial<T, dof> (basematerial.h:5-20)."""

    def __init__(self, type, E, rho, nu, name="Material"):
        self.type, self.E, self.rho, self.nu, self.name = type, E, rho, nu, name


class Elastoplastic(BaseMaterial):
    """Elastoplastic(E, rho, nu, beta, H, Y0, name) (elastoplastic.h:9-22).
    beta is kept and not used, as in the reference; hardening is linear
    isotropic with modulus H from the initial yield stress Y0."""

    def __init__(self, E, rho, nu, beta, H, Y0, name="Material"):
        super().__init__("InfinitesimalElastoplastic", E, rho, nu, name)
        self.beta, self.H, self.Y0 = beta, H, Y0


class Wall:
    """Wall<T, dim, Basis>(name, location, stiffness, slave_node_indices,
    num_slave_nodes, normal) (wall.h:13-44)."""

    def __init__(self, name, location, stiffness, slave_node_indices=None, num_slave_nodes=0,
                 normal=1, dim=2):
        if normal not in (1, -1):
            raise ValueError("normal must be +1 or -1")
        self.name, self.location, self.stiffness = name, location, stiffness
        self.slave_node_indices = slave_node_indices
        self.num_slave_nodes = num_slave_nodes
        self.normal, self.dim = normal, dim
        self.norm_stiffness = normal * stiffness


def loop_iterations(dt, time_end):
    """Iterations of `while (time <= time_end) { ...; time += dt; }`
    (dynamics.h:471-511) with the same floating-point accumulation."""
    t, n = 0.0, 0
    while t <= time_end:
        n += 1
        t += dt
    return n


class Dynamics:
    """Dynamics<T, Basis, Analysis, Quadrature> (dynamics.h:16-513) on the GPU.

    Options (keyword arguments) select the force model and kernels; the
    defaults reproduce the reference's step (compat physics, 5-point rule,
    row-sum mass, FP64).
    """

    def __init__(self, mesh, material, wall=None, **options):
        self.mesh, self.material, self.wall = mesh, material, wall
        self.ndof = mesh.num_nodes * 3
        self.options = options
        self.ctx = None
        n = mesh.num_nodes
        self.global_xloc = np.array(mesh.xloc, dtype=np.float64, copy=True)
        self.vel = np.zeros((n, 3))
        self.global_dof = np.zeros((n, 3))
        self.global_acc = np.zeros((n, 3))
        self.global_mass = np.zeros((n, 3))
        self.vel_i = np.zeros((n, 3))
        self.global_strains = np.zeros((n, 6))
        self.global_stress = np.zeros((n, 6))
        self.timestep = 0
        self._x0 = np.array(mesh.xloc, dtype=np.float64, copy=True)

    # -- reference API ----------------------------------------------------
    def initialize(self, init_position, init_velocity):
        """dynamics.h:99-111: shift the mesh, give every node init_velocity."""
        self.mesh.xloc = self.mesh.xloc + np.asarray(init_position, dtype=np.float64)
        self._x0 = np.array(self.mesh.xloc, copy=True)
        self.vel[:] = np.asarray(init_velocity, dtype=np.float64)

    def solve(self, dt, time_end, export_interval=2**31 - 1, on_export=None):
        """dynamics.h:414-512: init update + half-kick, then the loop until
        the accumulated time passes time_end.  The host arrays hold the final
        state on return.  on_export(self) is called on export steps with the
        host arrays refreshed (timestep % export_interval == 0)."""
        self.init(dt)
        n = loop_iterations(dt, time_end)
        self.timestep = 0
        done = 0
        while done < n:
            if export_interval > 0 and on_export is not None:
                nxt = ((self.timestep // export_interval) + 1) * export_interval
                chunk = min(n - done, nxt - self.timestep if self.timestep % export_interval
                            else 1)
            else:
                chunk = n - done
            self.step(chunk)
            done += chunk
            self.timestep += chunk
            if on_export is not None and (self.timestep - 1) % export_interval == 0:
                self.fetch()
                on_export(self)
        self.fetch()

    # -- extended API -----------------------------------------------------
    def _create(self):
        if self.ctx is not None:
            return
        m, mat, w = self.mesh, self.material, self.wall
        slave = getattr(m, "slave_nodes", None)
        desc = m._desc(slave if slave is not None and len(slave) else None)
        material = L.Material(mat.E, mat.rho, mat.nu, getattr(mat, "beta", 0.0),
                              getattr(mat, "H", 0.0), getattr(mat, "Y0", 0.0))
        wd = None
        if w is not None:
            wd = L.WallDesc(1, w.dim, w.normal, w.location, w.stiffness)
        opts = L.default_options(**self.options)
        ctx = C.c_void_p()
        L.check(L.lib().cpi_create(C.byref(desc), C.byref(material),
                                   C.byref(wd) if wd is not None else None, C.byref(opts),
                                   C.byref(ctx)))
        self.ctx = ctx

    def init(self, dt):
        """First half of solve(): mass, initial update and half-kick.
        dt <= 0 picks -dt (0.5 when 0) times the stable step."""
        self._create()
        x0 = np.ascontiguousarray(self._x0, dtype=np.float64)
        v0 = np.ascontiguousarray(self.vel, dtype=np.float64)
        L.check(L.lib().cpi_set_state(self.ctx, L.dptr(x0), L.dptr(v0)), self.ctx)
        L.check(L.lib().cpi_init(self.ctx, float(dt)), self.ctx)
        self.dt = self.info().dt

    def step(self, n):
        L.check(L.lib().cpi_step(self.ctx, int(n)), self.ctx)

    def synchronize(self):
        L.check(L.lib().cpi_synchronize(self.ctx), self.ctx)

    def fetch(self):
        """Downloads x, v (v_{n+1/2}), a_n, mass and v_n into the host arrays."""
        n = self.mesh.num_nodes
        x, v, a, vi = (np.zeros((n, 3)) for _ in range(4))
        m = np.zeros(n)
        L.check(L.lib().cpi_get_state(self.ctx, L.dptr(x), L.dptr(v), L.dptr(a), L.dptr(m),
                                      L.dptr(vi)), self.ctx)
        self.global_xloc, self.vel, self.global_acc, self.vel_i = x, v, a, vi
        self.global_mass = np.repeat(m[:, None], 3, axis=1)
        return x, v, a, m, vi

    def stress_strain(self):
        n = self.mesh.num_nodes
        st, ss = np.zeros((n, 6)), np.zeros((n, 6))
        L.check(L.lib().cpi_get_stress_strain(self.ctx, L.dptr(st), L.dptr(ss)), self.ctx)
        self.global_strains, self.global_stress = st, ss
        return st, ss

    def element_state(self):
        inf = self.info()
        sig = np.zeros((inf.num_elements, inf.num_quadrature_pts, 6))
        ep = np.zeros((inf.num_elements, inf.num_quadrature_pts))
        L.check(L.lib().cpi_get_element_state(self.ctx, L.dptr(sig), L.dptr(ep)), self.ctx)
        return sig, ep

    def energy(self):
        row = np.zeros(L.NENERGY)
        st = C.c_int64()
        L.check(L.lib().cpi_get_energy(self.ctx, L.dptr(row), C.byref(st)), self.ctx)
        return row, st.value

    def energy_history(self, cap=4096):
        rows = np.zeros((cap, L.NENERGY))
        cnt = C.c_int32()
        L.check(L.lib().cpi_get_energy_history(self.ctx, L.dptr(rows), cap, C.byref(cnt)),
                self.ctx)
        return rows[:cnt.value]

    def info(self):
        inf = L.Info()
        self._create()
        L.check(L.lib().cpi_get_info(self.ctx, C.byref(inf)), self.ctx)
        return inf

    def coloring(self):
        col = np.zeros(self.mesh.num_elements, dtype=np.int32)
        n = C.c_int32()
        L.check(L.lib().cpi_get_coloring(self.ctx, L.iptr(col), C.byref(n)), self.ctx)
        return col, n.value

    def stable_dt(self):
        self._create()
        a, b = C.c_double(), C.c_double()
        L.check(L.lib().cpi_stable_dt(self.ctx, C.byref(a), C.byref(b)), self.ctx)
        return a.value, b.value

    def stable_dt_noalloc(self):
        """stable_dt of the current configuration without any allocation
        (cpi_stable_dt_noalloc; needs init)."""
        self._create()
        a, b = C.c_double(), C.c_double()
        L.check(L.lib().cpi_stable_dt_noalloc(self.ctx, C.byref(a), C.byref(b)), self.ctx)
        return a.value, b.value

    def time_steps(self, n):
        ms = C.c_double()
        L.check(L.lib().cpi_time_steps(self.ctx, int(n), C.byref(ms)), self.ctx)
        return ms.value

    def profile_kernels(self, n):
        out = (L.KernelTime * 8)()
        cnt = C.c_int32()
        L.check(L.lib().cpi_profile_kernels(self.ctx, int(n), out, 8, C.byref(cnt)), self.ctx)
        return {out[i].name.decode(): (out[i].total_ms, out[i].launches) for i in range(cnt.value)}

    def debug_strain(self, alpha, def_case):
        """dynamics.h:303-412 on the GPU: (total strain energy, total volume)."""
        self._create()
        W, V = C.c_double(), C.c_double()
        L.check(L.lib().cpi_debug_strain(self.ctx, alpha, def_case, C.byref(W), C.byref(V)),
                self.ctx)
        return W.value, V.value

    # -- surrogate return map (stress_update=RETURN_MLP) -------------------
    def set_surrogate(self, mlp):
        """Installs a copy of a trained Surrogate's weights (before init)."""
        self._create()
        L.check(L.lib().cpi_set_surrogate(self.ctx, mlp.handle), self.ctx)

    def surrogate_stats(self):
        """(points that took the surrogate, points that fell back)."""
        a, f = C.c_int64(), C.c_int64()
        L.check(L.lib().cpi_get_surrogate_stats(self.ctx, C.byref(a), C.byref(f)), self.ctx)
        return a.value, f.value

    def reset_surrogate_stats(self):
        L.check(L.lib().cpi_reset_surrogate_stats(self.ctx), self.ctx)

    def enable_samples(self, capacity):
        """Records the exact return map's yielding points (physics EP)."""
        self._create()
        L.check(L.lib().cpi_samples_enable(self.ctx, int(capacity)), self.ctx)
        self._sample_cap = int(capacity)

    def subsample_samples(self, keep, seed=0):
        """Records only the yielding points whose hash of (seed, step,
        element, point) falls in the fraction keep (see
        cpi_samples_subsample)."""
        self._create()
        L.check(L.lib().cpi_samples_subsample(self.ctx, float(keep), int(seed)), self.ctx)

    def samples(self):
        """(recorded samples [k][20] float32, number of yielding points seen)."""
        cap = getattr(self, "_sample_cap", 0)
        out = np.zeros((max(cap, 1), L.SAMPLE_FLOATS), dtype=np.float32)
        cnt = C.c_int32()
        L.check(L.lib().cpi_samples_get(self.ctx, L.fptr(out), cap, C.byref(cnt)), self.ctx)
        return out[:min(cnt.value, cap)].copy(), cnt.value

    def reset_samples(self):
        L.check(L.lib().cpi_samples_reset(self.ctx), self.ctx)

    def close(self):
        if self.ctx is not None:
            L.lib().cpi_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
