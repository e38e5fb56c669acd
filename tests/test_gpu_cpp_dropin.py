"""The C++ drop-in layer (cppimpact_amd/host, driven by the example
programs that mirror the reference's test.cu and FEA_debug.cu) on the GPU:
FEA_debug's known answers and the reference driver's fuselage impact against
the CPU oracle."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import cppimpact_amd as ci
from oracle import pyoracle as O

from common import E, rel, run_oracle

pytestmark = pytest.mark.gpu
EX = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cppimpact_amd",
                  "examples")


def run(args, cwd):
    p = subprocess.run(args, cwd=cwd, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    return p.stdout


@pytest.mark.parametrize("case", range(6))
def test_fea_debug_driver_known_answers(golden, tmp_path, case):
    ref = json.load(open(os.path.join(golden, "fea_debug_values.json")))["cases"][case]
    exe = os.path.join(EX, f"fea_debug_{ref['basis']}")
    out = run([exe, os.path.join(golden, ref["mesh"]), str(ref["def"]), str(tmp_path)], tmp_path)
    W = float(re.search(r"Total Strain Energy = (\S+)", out).group(1))
    V = float(re.search(r"Total Volume = (\S+)", out).group(1))
    # the reference prints %f: compare to its printed digits
    assert abs(W - ref["W"]) <= 1e-6 * abs(ref["W"]) + 1e-6
    assert abs(V - ref["V"]) <= 1e-6 * abs(ref["V"]) + 1e-6
    assert (tmp_path / "simulation_0.vtk").exists()


def test_impact_driver_fuselage_against_oracle(golden, tmp_path):
    """test.cu's set-up (wall at 0.15 - 5e-5, k = 10 E, v0 = -10, dt = 1e-7,
    --smoke: t_end = 10 dt) on the 2647-element fuselage mesh."""
    mesh_path = os.path.join(golden, "fuselage_2647.inp")
    out = run([os.path.join(EX, "impact_linear"), "--smoke", "--mesh", mesh_path, "--dump",
               str(tmp_path / "s"), "--out", str(tmp_path)], tmp_path)
    steps = int(re.search(r"Steps: (\d+)", out).group(1))
    dt = 1e-7
    assert steps == ci.loop_iterations(dt, 10 * dt)
    m = ci.Mesh(4)
    assert m.load_mesh(mesh_path) == 0
    x0 = m.xloc
    xg = np.loadtxt(tmp_path / "s_x.txt").reshape(-1, 3)
    vg = np.loadtxt(tmp_path / "s_v.txt").reshape(-1, 3)
    mg = np.loadtxt(tmp_path / "s_m.txt").reshape(-1, 3)
    o = run_oracle(x0, m.element_nodes, (0, 0, -10.0), dt, steps, O.COMPAT, 10 * E,
                   wall_location=0.15 - 0.00005)
    assert rel(xg - x0, o["x"] - x0) < 1e-10
    assert rel(vg, o["v"]) < 1e-10
    assert rel(mg[:, 0], o["m"]) < 1e-12
    # the first loop iteration always exports (timestep 0 % INT_MAX == 0)
    assert (tmp_path / "simulation_0.vtk").exists()
    vtk = (tmp_path / "simulation_0.vtk").read_text().splitlines()
    assert vtk[0] == "# vtk DataFile Version 3.0" and vtk[4] == f"POINTS {m.num_nodes} float"


VTK_ORDER = ["POINTS", "CELLS", "CELL_TYPES", "POINT_DATA", "velocity", "strain1", "strain2",
             "stress1", "stress2", "acceleration", "mass"]


def parse_vtk(path, npe):
    """Sections of a legacy ASCII VTK file as export_to_vtk writes it
    (dynamics.h:113-254), in file order."""
    lines = path.read_text().splitlines()
    assert lines[:4] == ["# vtk DataFile Version 3.0", "FEA simulation data", "ASCII",
                         "DATASET UNSTRUCTURED_GRID"]
    order, data, i, n = [], {}, 4, None
    while i < len(lines):
        head = lines[i].split()
        if head[0] == "POINTS":
            n = int(head[1])
            assert head[2] == "float"
            key, rows = "POINTS", lines[i + 1:i + 1 + n]
        elif head[0] == "CELLS":
            ne = int(head[1])
            assert int(head[2]) == ne * (npe + 1)
            key, rows = "CELLS", lines[i + 1:i + 1 + ne]
        elif head[0] == "CELL_TYPES":
            key, rows = "CELL_TYPES", lines[i + 1:i + 1 + int(head[1])]
        elif head[0] == "POINT_DATA":
            assert int(head[1]) == n
            key, rows = "POINT_DATA", []
        else:
            assert head[0] == "VECTORS" and head[2] == "double", lines[i]
            key, rows = head[1], lines[i + 1:i + 1 + n]
        order.append(key)
        data[key] = np.array([[float(t) for t in r.split()] for r in rows]) if rows else None
        i += 1 + len(rows)
    return order, data


def test_vtk_export_and_probe_content(golden, tmp_path):
    """The fuselage pressed 0.2 mm into the wall from the start, exported
    every 10 loop iterations: the export of iteration 10 has the field order
    of dynamics.h:113-254 and every value equal to the oracle's state after
    11 iterations at the six decimals written (positions, v_n, the nodal
    strain and stress of update(), a_n, the mass on three columns), the
    connectivity and cell type 10.  The node probe (dynamics.h:256-274)
    starts afresh at iteration 0 and logs iterations 0 and 10 with the
    simulated time of the state it prints."""
    mesh_path = os.path.join(golden, "fuselage_2647.inp")
    run_dir = tmp_path / "run"
    run_dir.mkdir()
    (tmp_path / "output" / "nodes").mkdir(parents=True)
    (tmp_path / "output" / "nodes" / "node_96.txt").write_text("stale line from an earlier run\n")
    dt, wall = 1e-7, 0.1502
    run([os.path.join(EX, "impact_linear"), "--mesh", mesh_path, "--wall", str(wall), "--tend",
         str(15 * dt), "--export-interval", "10", "--out", str(tmp_path)], run_dir)
    m = ci.Mesh(4)
    assert m.load_mesh(mesh_path) == 0
    order, d = parse_vtk(tmp_path / "simulation_10.vtk", 4)
    assert order == VTK_ORDER
    s = O.Solver(m.xloc, m.element_nodes, (0, 0, -10.0), E=E, nu=0.33, rho=2700.0,
                 physics=O.COMPAT, wall=dict(location=wall, stiffness=10 * E), dt=dt)
    o = s.solve(11, with_strain=True)
    half = 5.01e-7  # half a unit of the sixth decimal

    def close(a, b):
        return np.abs(a - b).max() <= half + 1e-9 * np.abs(b).max()

    assert close(d["POINTS"], o["x"])
    assert close(d["velocity"], o["v_i"])
    assert close(d["acceleration"], o["a"])
    assert close(d["mass"], np.repeat(o["m"][:, None], 3, axis=1))
    assert close(np.hstack([d["strain1"], d["strain2"]]), o["strain"])
    assert close(np.hstack([d["stress1"], d["stress2"]]), o["stress"])
    # the contact has loaded the mesh: stresses far above the written
    # resolution, so the comparison above has teeth
    assert np.abs(o["stress"]).max() > 1e3
    assert np.array_equal(d["CELLS"][:, 0], np.full(m.num_elements, 4.0))
    assert np.array_equal(d["CELLS"][:, 1:].astype(np.int64), m.element_nodes)
    assert np.all(d["CELL_TYPES"] == 10)
    probe = (tmp_path / "output" / "nodes" / "node_96.txt").read_text().splitlines()
    heads = [ln for ln in probe if ln.startswith("Timestep")]
    assert heads == [f"Timestep 0, Time: {dt:.2f}s:", f"Timestep 10, Time: {11 * dt:.2f}s:"]
    i10 = probe.index(heads[1])
    pos = [float(t) for t in probe[i10 + 1].split("(")[1].rstrip(")").split(",")]
    assert np.allclose(pos, o["x"][96], rtol=1e-5, atol=0)


def test_vtk_export_tet10_cell_type(tmp_path):
    """A tet10 mesh exports ten node ids per cell and VTK cell type 24
    (quadratic tetra)."""
    x, conn = ci.box_mesh(2, 2, 2, 0.1, 0.1, 0.1, origin=(0.0, 0.0, 0.2), quadratic=True)
    m = ci.Mesh.from_arrays(x, conn)
    path = str(tmp_path / "box10.inp")
    m.write_inp(path)
    run([os.path.join(EX, "impact_quadratic"), "--smoke", "--mesh", path, "--wall", "-1",
         "--out", str(tmp_path)], tmp_path)
    order, d = parse_vtk(tmp_path / "simulation_0.vtk", 10)
    assert order == VTK_ORDER
    assert np.all(d["CELL_TYPES"] == 24) and len(d["CELL_TYPES"]) == conn.shape[0]
    assert np.array_equal(d["CELLS"][:, 0], np.full(conn.shape[0], 10.0))
    assert np.array_equal(d["CELLS"][:, 1:].astype(np.int64), conn)
