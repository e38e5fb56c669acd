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
