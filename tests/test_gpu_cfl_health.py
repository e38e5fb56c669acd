"""The stable time step and the numerical-health paths on the GPU.

cpi_stable_dt (element CFL limit and wall-spring limit 2 sqrt(m_min / k))
and the per-step stable-dt row of the energy reduction against the oracle's
orc_stable_dt / ORC_E_MINDT on c0, c2 and the tet10 c1 mesh; the dt cpi_init
selects; and the conditions that must end in CPI_ERR_NUMERIC (a dt above
the stable limit, an element inverted by the motion, a state that
overflows) or be refused up front (a NaN initial velocity).  The reference
steps with a fixed dt and checks none of this (test.cu:83; analysis.h:235-239
only prints a non-positive detJ)."""
import numpy as np
import pytest

import cppimpact_amd as ci
from cppimpact_amd import _lib as L
from cppimpact_amd import workloads as W
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def oracle_of(w, speed=None, dt=1e-9):
    v = w.speed if speed is None else speed
    return O.Solver(w.x, w.conn, np.asarray(w.direction) * v, E=W.E, nu=W.NU, rho=W.RHO,
                    sigma_y=W.SY, H=W.H, physics=O.EP,
                    wall=dict(location=0.0, stiffness=w.wall_stiffness), dt=dt)


@pytest.mark.parametrize("name", ["c0", "c1", "c2"])
def test_stable_dt_and_selected_dt_match_oracle(name):
    w = getattr(W, name)()
    d = w.dynamics()
    dte, dtc = d.stable_dt()
    ode, odc = oracle_of(w).stable_dt()
    assert abs(dte - ode) <= 1e-12 * ode and abs(dtc - odc) <= 1e-12 * odc
    # closed form for the undistorted cube (c0): every tet of the 6-tet
    # split shares a main diagonal, with volume a^3/6 and largest face a^2 /
    # sqrt(2), so its smallest altitude 3V/A is a / sqrt(2); the limit is
    # that over the tet4 length factor 2 and the wave speed sqrt(c11 / rho)
    if name == "c0":
        c11 = W.E * (1 - W.NU) / ((1 + W.NU) * (1 - 2 * W.NU))
        a = W.SIDE / 10
        assert abs(dte - a / np.sqrt(2.0) / 2.0 / np.sqrt(c11 / W.RHO)) <= 1e-12 * dte
    d.init(0.0)
    assert abs(d.info().dt - 0.5 / np.sqrt(1 / ode ** 2 + 1 / odc ** 2)) <= 1e-12 * d.info().dt
    d.close()


@pytest.mark.parametrize("name,steps", [("c0", 40), ("c1", 3), ("c2", 3)])
def test_step_stable_dt_row_matches_oracle(name, steps):
    """E_MINDT of the last step: the element CFL limit of the configuration
    that step started from, reduced over waves on the device."""
    w = getattr(W, name)()
    d = w.dynamics()
    d.init(0.0)
    dt = d.info().dt
    d.step(steps)
    e, _ = d.energy()
    o = oracle_of(w, dt=dt).solve(steps)
    om = o["energy"][-1][O.E_MINDT]
    assert abs(e[L.E_MINDT] - om) <= 1e-12 * om
    d.close()


def test_dt_above_stable_limit_is_reported():
    w = W.c0()
    d = w.dynamics()
    dte, dtc = d.stable_dt()
    d.init(1.05 * min(dte, dtc))
    d.step(1)
    with pytest.raises(L.CpiError) as ex:
        d.synchronize()
    assert ex.value.code == -5 and "stable limit" in str(ex.value)
    # half the limit steps quietly
    d.init(0.5 * min(dte, dtc))
    d.step(5)
    d.synchronize()
    d.close()


def test_element_inverted_by_the_motion_is_reported():
    """One node thrown through its elements in the first step: detJ <= 0."""
    w = W.c0()
    d = w.dynamics()
    d.vel[:] = 0.0
    d.vel[w.x.shape[0] // 2] = (0.0, 0.0, 2e5)
    d.init(0.0)
    # the first step moves the node, the second computes its elements from
    # the moved positions
    d.step(2)
    with pytest.raises(L.CpiError) as ex:
        d.synchronize()
    assert ex.value.code == -5 and "inverted" in str(ex.value)
    d.close()


def test_overflowing_state_is_reported():
    """The reference's force model at forty times its stable step diverges
    until the energies overflow."""
    w = W.c0()
    d = ci.Dynamics(ci.Mesh.from_arrays(w.x, w.conn), W.material(),
                    ci.Wall("Wall", 0.0, w.wall_stiffness, None, 0, 1))
    d.vel[:] = (0.0, 0.0, -10.0)
    dte, dtc = d.stable_dt()
    d.init(40 * min(dte, dtc))
    with pytest.raises(L.CpiError) as ex:
        for _ in range(100):
            d.step(10)
            d.synchronize()
    assert ex.value.code == -5
    d.close()


def test_nan_initial_velocity_is_refused():
    w = W.c0()
    d = w.dynamics()
    d.vel[3] = (np.nan, 0.0, 0.0)
    with pytest.raises(L.CpiError) as ex:
        d.init(0.0)
    assert ex.value.code == -1 and "non-finite" in str(ex.value)
    d.close()
