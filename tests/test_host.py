"""Host-side logic of libcppimpact_hip.so on the CPU: the C ABI surface,
mesh generators, the .inp reader/writer, greedy colouring (bit-exact with
the oracle) and the error path without a GPU."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

import cppimpact_amd as ci
from cppimpact_amd import _lib as L
from oracle import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "cppimpact_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(cpi_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 25
    lib = C.CDLL(L.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    # and the Python bindings cover exactly the declared set
    assert sorted(L.SIGNATURES) == names


def signed_volumes(x, conn):
    o = 0 if conn.shape[1] == 4 else 2
    i1, i2 = (1, 2) if conn.shape[1] == 4 else (0, 1)
    a = x[conn[:, i1]] - x[conn[:, o]]
    b = x[conn[:, i2]] - x[conn[:, o]]
    d = x[conn[:, 3]] - x[conn[:, o]]
    return np.einsum("ij,ij->i", a, np.cross(b, d)) / 6


def face_counts(conn):
    faces = np.sort(np.concatenate([conn[:, [1, 2, 3]], conn[:, [0, 2, 3]], conn[:, [0, 1, 3]],
                                    conn[:, [0, 1, 2]]]), axis=1)
    _, counts = np.unique(faces, axis=0, return_counts=True)
    return counts


@pytest.mark.parametrize("quadratic", [False, True])
def test_box_mesh_counts_volume_orientation_conformity(quadratic):
    n = (3, 4, 5)
    x, conn = ci.box_mesh(*n, 0.3, 0.4, 0.5, origin=(1.0, -2.0, 0.5), quadratic=quadratic)
    if quadratic:
        assert x.shape[0] == (2 * n[0] + 1) * (2 * n[1] + 1) * (2 * n[2] + 1)
        assert conn.shape == (6 * n[0] * n[1] * n[2], 10)
        # mid-side nodes sit exactly at edge midpoints
        for q, (a, b) in enumerate([(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]):
            assert np.allclose(x[conn[:, 4 + q]], 0.5 * (x[conn[:, a]] + x[conn[:, b]]),
                               atol=1e-15)
    else:
        assert x.shape[0] == (n[0] + 1) * (n[1] + 1) * (n[2] + 1)
        assert conn.shape == (6 * n[0] * n[1] * n[2], 4)
    v = signed_volumes(x, conn)
    assert v.min() > 0
    assert abs(v.sum() - 0.3 * 0.4 * 0.5) < 1e-14
    # conforming: every face is shared by at most two tets, and the boundary
    # faces (count 1) tile the surface: 2 per cell face
    counts = face_counts(conn[:, :4])
    assert counts.max() == 2
    nb = 2 * 2 * (n[0] * n[1] + n[1] * n[2] + n[0] * n[2])
    assert (counts == 1).sum() == nb
    assert np.allclose(x.min(axis=0), (1.0, -2.0, 0.5)) and np.allclose(
        x.max(axis=0), (1.3, -1.6, 1.0))


def test_box_jitter_is_seeded_interior_only_and_bounded():
    n, h = 6, 0.25 / 6
    x0, c0 = ci.box_mesh(n, n, n, 0.25, 0.25, 0.25)
    x1, c1 = ci.box_mesh(n, n, n, 0.25, 0.25, 0.25, jitter=0.1, seed=42)
    x2, _ = ci.box_mesh(n, n, n, 0.25, 0.25, 0.25, jitter=0.1, seed=42)
    x3, _ = ci.box_mesh(n, n, n, 0.25, 0.25, 0.25, jitter=0.1, seed=43)
    assert np.array_equal(x1, x2) and not np.array_equal(x1, x3)
    d = np.abs(x1 - x0)
    boundary = np.any((x0 < 1e-12) | (x0 > 0.25 - 1e-12), axis=1)
    assert d[boundary].max() == 0.0
    assert 0.05 * h < d[~boundary].max() <= 0.1 * h + 1e-17
    assert np.array_equal(c0, c1)
    assert signed_volumes(x1, c1).min() > 0


def test_c2_and_c1_sizes():
    # the 1.57M-tet cube and the tet10 cube of the configurations
    nn, ne = C.c_int32(), C.c_int32()
    s = L.BoxSpec(64, 64, 64, 0, 0.25, 0.25, 0.25, (C.c_double * 3)(0, 0, 0), 0.1, 0)
    L.check(L.lib().cpi_mesh_box(C.byref(s), C.byref(nn), C.byref(ne), None, None))
    assert (nn.value, ne.value) == (274625, 1572864)
    s = L.BoxSpec(20, 20, 20, 1, 0.25, 0.25, 0.25, (C.c_double * 3)(0, 0, 0), 0.0, 0)
    L.check(L.lib().cpi_mesh_box(C.byref(s), C.byref(nn), C.byref(ne), None, None))
    assert (nn.value, ne.value) == (68921, 48000)
    s = L.CylSpec(360, 185, 3, 1, 1.0, 4.0, 0.02, (C.c_double * 3)(0, 0, 0))
    L.check(L.lib().cpi_mesh_cylinder(C.byref(s), C.byref(nn), C.byref(ne), None, None))
    assert (nn.value, ne.value) == (720 * 371 * 7, 360 * 185 * 3 * 6)


@pytest.mark.parametrize("quadratic", [False, True])
def test_cylinder_periodic_seam_and_volume(quadratic):
    nt, nz, nr, R, Lz, t = 48, 6, 2, 1.0, 0.5, 0.05
    x, conn = ci.cylinder_mesh(nt, nz, nr, R, Lz, t, quadratic=quadratic)
    npts = (2 * nt) * (2 * nz + 1) * (2 * nr + 1) if quadratic else nt * (nz + 1) * (nr + 1)
    assert x.shape[0] == npts
    v = signed_volumes(x, conn)
    assert v.min() > 0
    # straight-chord prisms: annulus volume times sin(2pi/n)/(2pi/n)
    th = 2 * math.pi / nt
    exact = math.pi * ((R + t / 2) ** 2 - (R - t / 2) ** 2) * Lz * math.sin(th) / th
    assert abs(v.sum() - exact) < 1e-12
    counts = face_counts(conn[:, :4])
    assert counts.max() == 2
    # closed around theta: only the two end annuli and the two cylindrical
    # skins are boundary, 2 triangles per quad
    assert (counts == 1).sum() == 2 * (2 * nt * nr + 2 * nt * nz)
    # corners lie on the radii; mid-sides are chord midpoints, just inside
    corners = np.unique(conn[:, :4])
    r = np.hypot(x[corners, 0], x[corners, 1])
    assert r.min() >= R - t / 2 - 1e-12 and r.max() <= R + t / 2 + 1e-12


@pytest.mark.parametrize("case", ["c0", "jitter", "tet10", "fuselage"])
def test_colouring_bit_exact_with_oracle(case, golden):
    if case == "c0":
        x, conn = ci.box_mesh(10, 10, 10, 0.25, 0.25, 0.25)
    elif case == "jitter":
        x, conn = ci.box_mesh(16, 12, 9, 0.25, 0.2, 0.1, jitter=0.2, seed=3)
    elif case == "tet10":
        x, conn = ci.box_mesh(6, 6, 6, 0.25, 0.25, 0.25, quadratic=True)
    else:
        m = ci.Mesh(4)
        assert m.load_mesh(os.path.join(golden, "fuselage_2647.inp")) == 0
        x, conn = m.xloc, m.element_nodes
    col, n = ci.greedy_coloring(conn, x.shape[0])
    ref, nref = O.greedy_coloring(conn, x.shape[0])
    assert n == nref and np.array_equal(col, ref)


def test_inp_reader_on_reference_fixtures(golden):
    m = ci.Mesh(4)
    assert m.load_mesh(os.path.join(golden, "linear_tet.inp")) == 0
    assert (m.num_nodes, m.num_elements) == (4, 1)
    assert np.allclose(m.xloc, [[0, 0, 0], [10, 0, 0], [0, 10, 0], [0, 0, 10]])
    assert m.element_nodes.tolist() == [[0, 1, 2, 3]]
    assert m.node_set_names == ["slave"] and m.slave_nodes.tolist() == [0]
    q = ci.Mesh(10)
    assert q.load_mesh(os.path.join(golden, "quad_tet.inp")) == 0  # trailing comma on node 3
    assert (q.num_nodes, q.num_elements) == (10, 1)
    assert np.allclose(q.xloc[2], [0, 0, 0]) and np.allclose(q.xloc[4], [5, 5, 0])
    f = ci.Mesh(4)
    assert f.load_mesh(os.path.join(golden, "fuselage_2647.inp")) == 0
    assert (f.num_nodes, f.num_elements) == (1099, 2647)
    assert f.num_slave_nodes == f.num_nodes  # the fuselage slave set covers every node
    c = ci.Mesh(4)
    assert c.load_mesh(os.path.join(golden, "cube_linear_5758.inp")) == 0
    assert (c.num_nodes, c.num_elements) == (1443, 5758)
    assert 0 < c.num_slave_nodes < c.num_nodes
    assert signed_volumes(c.xloc, c.element_nodes).min() > 0


def test_inp_missing_file_returns_one_like_the_reference(tmp_path):
    m = ci.Mesh(4)
    assert m.load_mesh(str(tmp_path / "nope.inp")) == 1


def test_inp_write_read_round_trip(tmp_path):
    x, conn = ci.box_mesh(3, 2, 2, 0.3, 0.2, 0.2, quadratic=True, jitter=0.1, seed=1)
    m = ci.Mesh.from_arrays(x, conn, slave_nodes=np.arange(0, x.shape[0], 7, dtype=np.int32),
                            fixed_nodes=np.arange(1, x.shape[0], 5, dtype=np.int32))
    p = str(tmp_path / "m.inp")
    m.write_inp(p)
    r = ci.Mesh(10)
    assert r.load_mesh(p) == 0
    assert np.array_equal(r.xloc, x)  # %.17g round-trips doubles
    assert np.array_equal(r.element_nodes, conn)
    assert np.array_equal(r.slave_nodes, m.slave_nodes)
    assert np.array_equal(r.fixed_nodes, m.fixed_nodes)
    assert r.node_set_names == ["fixed", "slave"]


def test_loop_iteration_count_follows_float_accumulation():
    # while (time <= time_end) { ...; time += dt; }  (dynamics.h:471-511)
    for dt, t_end in ((1e-7, 1e-6), (1e-7, 10 * 1e-7), (0.1, 1.0), (2.5e-7, 1e-5)):
        t, n = 0.0, 0
        while t <= t_end:
            n += 1
            t += dt
        assert ci.loop_iterations(dt, t_end) == n
    assert ci.loop_iterations(1e-7, 1e-7 * 10) == 11


def test_create_without_gpu_fails_with_a_status_not_a_crash():
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    x, conn = ci.box_mesh(2, 2, 2, 1, 1, 1)
    dyn = ci.Dynamics(ci.Mesh.from_arrays(x, conn), ci.Elastoplastic(1e9, 1e3, 0.3, 1, 1, 1))
    with pytest.raises(ci.CpiError) as ei:
        dyn.init(1e-7)
    assert ei.value.code == -8  # CPI_ERR_NO_DEVICE


def test_invalid_inputs_are_rejected_before_the_device():
    lib = L.lib()
    x, conn = ci.box_mesh(2, 2, 2, 1, 1, 1)
    bad = conn.copy()
    bad[0, 0] = 10**6
    col = np.zeros(len(bad), np.int32)
    assert lib.cpi_greedy_coloring(len(bad), 4, L.iptr(bad), x.shape[0], L.iptr(col)) == -4
    s = L.BoxSpec(0, 2, 2, 0, 1, 1, 1, (C.c_double * 3)(0, 0, 0), 0.0, 0)
    nn, ne = C.c_int32(), C.c_int32()
    assert lib.cpi_mesh_box(C.byref(s), C.byref(nn), C.byref(ne), None, None) == -1
    assert b"invalid" in lib.cpi_last_error(None)


def test_cpp_driver_reports_missing_mesh():
    """Mesh::load_mesh returns 1 for a file it cannot open (mesh.h:57-70);
    the driver stops before any GPU call, as test.cu:67-71 does."""
    import subprocess
    exe = os.path.join(ROOT, "cppimpact_amd", "examples", "impact_linear")
    if not os.path.exists(exe):
        pytest.skip("example drivers not built")
    p = subprocess.run([exe, "--mesh", "/nonexistent/mesh.inp"], capture_output=True, text=True,
                       timeout=60)
    assert p.returncode == 1 and "Error loading mesh" in p.stderr


def test_create_refuses_an_inverted_element():
    """One element with two corners swapped has a negative volume: the
    context is refused with CPI_ERR_MESH naming it (the reference only
    prints a non-positive detJ when it meets one, analysis.h:235-239).  The
    mesh is validated before any device is touched."""
    import cppimpact_amd as ci
    from cppimpact_amd import _lib as L
    x, conn = ci.box_mesh(2, 2, 2, 0.1, 0.1, 0.1)
    conn = conn.copy()
    conn[5, [1, 2]] = conn[5, [2, 1]]
    d = ci.Dynamics(ci.Mesh.from_arrays(x, conn), ci.Elastoplastic(68.9e9, 2700.0, 0.33, 1.0,
                                                                    700e6, 276e6))
    with pytest.raises(L.CpiError) as ex:
        d.init(0.0)
    assert ex.value.code == -4 and "element 5 is inverted" in str(ex.value)


@pytest.mark.parametrize("what", ["no elements", "no nodes", "npe 7", "physics 5", "scatter 9",
                                  "E 0", "nu 0.5", "wall dim 3", "wall normal 0",
                                  "surrogate on tet10", "surrogate in FP64", "stream_state 3"])
def test_create_rejects_invalid_descriptions_before_the_device(what):
    """cpi_create validates the mesh, options, material and wall before it
    touches a device, so each bad description is CPI_ERR_INVALID with a
    message here too (the reference has no such checks)."""
    lib = L.lib()
    quad = what == "surrogate on tet10"
    x, conn = ci.box_mesh(2, 2, 2, 0.1, 0.1, 0.1, quadratic=quad)
    x = np.ascontiguousarray(x, np.float64)
    conn = np.ascontiguousarray(conn, np.int32)
    desc = L.MeshDesc(x.shape[0], conn.shape[0], conn.shape[1], L.dptr(x), L.iptr(conn), 0,
                      None, 0, None)
    mat = L.Material(68.9e9, 2700.0, 0.33, 1.0, 700e6, 276e6)
    wall = L.WallDesc(1, 2, 1, 0.0, 1e9)
    opts = {}
    if what == "no elements":
        desc.num_elements = 0
    elif what == "no nodes":
        desc.num_nodes = 0
    elif what == "npe 7":
        desc.nodes_per_element = 7
    elif what == "physics 5":
        opts["physics"] = 5
    elif what == "scatter 9":
        opts["scatter"] = 9
    elif what == "E 0":
        mat.E = 0.0
    elif what == "nu 0.5":
        mat.nu = 0.5
    elif what == "wall dim 3":
        wall.dim = 3
    elif what == "wall normal 0":
        wall.normal = 0
    elif what == "surrogate on tet10":
        opts.update(physics=L.PHYSICS_EP, precision=L.PREC_F32, stress_update=L.RETURN_MLP)
    elif what == "surrogate in FP64":
        opts.update(physics=L.PHYSICS_EP, stress_update=L.RETURN_MLP)
    elif what == "stream_state 3":
        opts["stream_state"] = 3
    o = L.default_options(**opts)
    ctx = C.c_void_p()
    rc = lib.cpi_create(C.byref(desc), C.byref(mat), C.byref(wall), C.byref(o), C.byref(ctx))
    assert rc == -1, (what, rc)
    assert lib.cpi_last_error(None), what
