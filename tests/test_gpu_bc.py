"""GPU parity of the node sets against the oracle: the wall on the slave set
only (cpi_options.contact_slave_only) and held nodes from the fixed set
(cpi_options.fixed_bc), both off by default as in the reference."""
import numpy as np
import pytest

from cppimpact_amd import _lib as L
from oracle import pyoracle as O

from common import cube, rel, run_gpu, run_oracle, wall_k

pytestmark = pytest.mark.gpu

SCATTERS = [L.SCATTER_BLOCK, L.SCATTER_COLOR, L.SCATTER_ATOMIC]


def _energies_match(g, o, tol):
    eo = o["energy"][-1]
    scale = max(abs(eo[O.E_KINETIC]), abs(eo[O.E_INTERNAL]))
    for k in (O.E_KINETIC, O.E_INTERNAL, O.E_PLASTIC, O.E_PENALTY, O.E_GRAVITY, O.E_CONTACTF):
        gk = g["energy"][k]
        assert abs(gk - eo[k]) <= tol * max(abs(eo[k]), 1e-3 * scale), (k, gk, eo[k])


@pytest.mark.parametrize("scatter", SCATTERS)
def test_held_nodes_tet4_match_oracle(scatter):
    x, conn = cube(4, gap=0.0, jitter=0.1, seed=3)
    fixed = np.where(x[:, 0] < 1e-9)[0]
    v0 = (0.0, 8.0, -3.0)
    g = run_gpu(x, conn, v0, 2e-7, 150, physics=L.PHYSICS_EP, scatter=scatter,
                fixed_nodes=fixed, fixed_bc=1)
    o = run_oracle(x, conn, v0, 2e-7, 150, O.EP, fixed_nodes=fixed)
    assert np.array_equal(g["x"][fixed], x[fixed])
    assert not g["v"][fixed].any()
    assert rel(g["x"] - x, o["x"] - x) < 1e-9
    assert rel(g["v"], o["v"]) < 1e-9
    _energies_match(g, o, 1e-8)
    g["dyn"].close()


def test_held_nodes_tet10_match_oracle():
    x, conn = cube(2, quadratic=True, gap=0.0)
    fixed = np.where(x[:, 2] < 1e-9)[0]
    v0 = (6.0, 0.0, 0.0)
    g = run_gpu(x, conn, v0, 5e-8, 80, physics=L.PHYSICS_EP, quadrature=L.QUAD_EXACT4,
                fixed_nodes=fixed, fixed_bc=1)
    o = run_oracle(x, conn, v0, 5e-8, 80, O.EP, fixed_nodes=fixed)
    assert np.array_equal(g["x"][fixed], x[fixed])
    assert rel(g["x"] - x, o["x"] - x) < 1e-9
    assert rel(g["v"], o["v"]) < 1e-9
    _energies_match(g, o, 1e-8)
    g["dyn"].close()


def test_fixed_set_ignored_unless_enabled():
    x, conn = cube(3, gap=0.0)
    fixed = np.where(x[:, 0] < 1e-9)[0]
    g = run_gpu(x, conn, (0.0, 8.0, 0.0), 2e-7, 40, physics=L.PHYSICS_EP, fixed_nodes=fixed)
    o = run_oracle(x, conn, (0.0, 8.0, 0.0), 2e-7, 40, O.EP)
    assert rel(g["x"] - x, o["x"] - x) < 1e-9
    g["dyn"].close()


@pytest.mark.parametrize("scatter", SCATTERS)
def test_slave_contact_matches_oracle(scatter):
    x, conn = cube(4, gap=1e-6)
    k = wall_k(0.25, 4)
    bottom = np.where(x[:, 2] < 1e-5)[0]
    sub = bottom[np.round(x[bottom, 0] / (0.25 / 4)).astype(int) % 2 == 0]
    v0 = (0.0, 0.0, -20.0)
    g = run_gpu(x, conn, v0, 2e-7, 100, k, physics=L.PHYSICS_EP, scatter=scatter,
                slave_nodes=sub, contact_slave_only=1)
    o = run_oracle(x, conn, v0, 2e-7, 100, O.EP, k, contact_nodes=sub)
    assert o["energy"][-1][O.E_CONTACTF] > 0
    assert rel(g["x"] - x, o["x"] - x) < 1e-9
    assert rel(g["v"], o["v"]) < 1e-9
    _energies_match(g, o, 1e-8)
    g["dyn"].close()
