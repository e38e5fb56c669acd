"""The oracle's threaded solve (the multi-core CPU baseline) against its
serial solve: the threaded element pass writes per-element contributions
and each node sums them in ascending element order, so every state array
and every energy row must be bit-identical, for any thread count, with
plastic flow, contact, held nodes and both element types."""
import numpy as np
import pytest

from oracle import pyoracle as O

from common import E, H, NU, RHO, SY, cube, wall_k

KEYS = ("x", "v", "a", "m", "v_i", "sig", "epbar", "energy")


def solver(x, conn, v0, physics, **kw):
    return O.Solver(x, conn, v0, E=E, nu=NU, rho=RHO, sigma_y=SY, H=H, physics=physics,
                    wall=dict(location=0.0, stiffness=wall_k(0.25, 3)), dt=2e-7, **kw)


def assert_identical(a, b):
    for k in KEYS:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("threads", [2, 3, 7])
def test_threaded_ep_tet4_plastic_impact_is_bit_identical(threads):
    x, conn = cube(3, jitter=0.1, seed=3, gap=1e-6)
    s = solver(x, conn, (0, 0, -40.0), O.EP)
    ref = s.solve(80)
    assert ref["epbar"].max() > 0.0 and ref["energy"][-1][O.E_CONTACTF] > 0.0
    assert_identical(s.solve(80, threads=threads), ref)


def test_threaded_ep_tet10_with_node_sets_is_bit_identical():
    x, conn = cube(2, quadratic=True, jitter=0.05, seed=5, gap=1e-6)
    top = np.where(x[:, 2] > x[:, 2].max() - 1e-9)[0]
    bottom = np.where(x[:, 2] < 1e-5)[0]
    s = solver(x, conn, (0, 0, -30.0), O.EP, fixed_nodes=top[::3], contact_nodes=bottom[::2])
    assert_identical(s.solve(60, threads=4), s.solve(60))


def test_threaded_compat_is_bit_identical():
    x, conn = cube(3, gap=1e-6)
    s = solver(x, conn, (0, 0, -10.0), O.COMPAT)
    assert_identical(s.solve(50, threads=5), s.solve(50))


def test_more_threads_than_elements():
    x, conn = cube(1, gap=1e-6)
    s = solver(x, conn, (0, 0, -10.0), O.EP)
    assert_identical(s.solve(20, threads=16), s.solve(20))


def test_strain_output_takes_the_serial_loop():
    x, conn = cube(2, gap=1e-6)
    s = solver(x, conn, (0, 0, -10.0), O.COMPAT)
    a, b = s.solve(10, with_strain=True, threads=4), s.solve(10, with_strain=True)
    assert_identical(a, b)
    assert np.array_equal(a["strain"], b["strain"]) and np.array_equal(a["stress"], b["stress"])
