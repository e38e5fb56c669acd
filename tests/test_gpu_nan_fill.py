"""Every device buffer filled with NaN at creation changes nothing.

With cpi_options.debug_fill_nan every allocation of the context starts as
all-ones bytes (NaN in any float or double).  A kernel that read memory the
context never wrote -- a partial-force slot no block fills, a stale energy
row, an unwritten queue row -- would carry the NaN into the state and the
energy check.  Each configuration of the parity tests runs with and without
the fill and must give the same state bit for bit (the atomic scatter, whose
summation order varies, to rounding), and match the oracle as before.  The
tet10 FP64 REF4 block kernel is the one that once gave non-finite energies
only after other kernels had run in the same process."""
import numpy as np
import pytest

import cppimpact_amd as ci
from cppimpact_amd import _lib as L
from oracle import pyoracle as O

from common import E, H, NU, RHO, SY, cube, rel, wall_k

pytestmark = pytest.mark.gpu

CASES = [
    # (id, quadratic, physics, precision, quadrature, scatter, v0, dt, steps)
    ("tet4-ep-f64-block", False, L.PHYSICS_EP, L.PREC_F64, L.QUAD_DEFAULT, L.SCATTER_BLOCK,
     30.0, 2e-7, 80),
    ("tet4-ep-f32-block", False, L.PHYSICS_EP, L.PREC_F32, L.QUAD_DEFAULT, L.SCATTER_BLOCK,
     30.0, 2e-7, 80),
    ("tet4-compat-f64-colour", False, L.PHYSICS_COMPAT, L.PREC_F64, L.QUAD_DEFAULT,
     L.SCATTER_COLOR, 10.0, 1e-7, 40),
    ("tet4-ep-f64-atomic", False, L.PHYSICS_EP, L.PREC_F64, L.QUAD_DEFAULT, L.SCATTER_ATOMIC,
     30.0, 2e-7, 80),
    ("tet10-ep-f64-ref4-block", True, L.PHYSICS_EP, L.PREC_F64, L.QUAD_REF4, L.SCATTER_BLOCK,
     30.0, 5e-8, 60),
    ("tet10-ep-f32-exact4-block", True, L.PHYSICS_EP, L.PREC_F32, L.QUAD_EXACT4,
     L.SCATTER_BLOCK, 30.0, 5e-8, 60),
    ("tet10-compat-f64-ref5-block", True, L.PHYSICS_COMPAT, L.PREC_F64, L.QUAD_REF5,
     L.SCATTER_BLOCK, 30.0, 2e-9, 4),
]


def run(x, conn, v0, dt, steps, **opts):
    d = ci.Dynamics(ci.Mesh.from_arrays(x, conn), ci.Elastoplastic(E, RHO, NU, 1.0, H, SY),
                    ci.Wall("Wall", 0.0, wall_k(0.25, 6), None, 0, 1), **opts)
    d.vel[:] = (0.0, 0.0, -v0)
    d.init(dt)
    d.step(steps)
    x1, v1, a1, m1, _ = d.fetch()
    sig, ep = d.element_state()
    e, _ = d.energy()
    d.close()
    return dict(x=x1, v=v1, a=a1, m=m1, sig=sig, ep=ep, e=e)


@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_nan_filled_buffers_change_nothing(case):
    _, quad, phys, prec, rule, scatter, v0, dt, steps = case
    x, conn = cube(4 if quad else 6, quadratic=quad, jitter=0.05, seed=11, gap=1e-6)
    opts = dict(physics=phys, precision=prec, quadrature=rule, scatter=scatter,
                block_elements=37 if quad else 64)
    plain = run(x, conn, v0, dt, steps, **opts)
    filled = run(x, conn, v0, dt, steps, debug_fill_nan=1, **opts)
    for k in ("x", "v", "a", "m", "sig", "ep", "e"):
        assert np.all(np.isfinite(filled[k])), k
        if scatter == L.SCATTER_ATOMIC:
            assert np.allclose(filled[k], plain[k], rtol=1e-10, atol=1e-12 * np.abs(plain[k]).max()), k
        else:
            assert np.array_equal(filled[k], plain[k]), k
    if prec == L.PREC_F64:
        o = O.Solver(x, conn, (0.0, 0.0, -v0), E=E, nu=NU, rho=RHO, sigma_y=SY, H=H,
                     physics=O.COMPAT if phys == L.PHYSICS_COMPAT else O.EP,
                     wall=dict(location=0.0, stiffness=wall_k(0.25, 6)), dt=dt,
                     quad_rule=None if rule == L.QUAD_DEFAULT else rule).solve(steps)
        assert rel(filled["x"] - x, o["x"] - x) < 1e-9
        assert rel(filled["v"], o["v"]) < 1e-9


def test_nan_filled_buffers_change_nothing_with_the_surrogate():
    """The surrogate step's queue, block list and counters too: a plastic
    impact with a trained network, with and without the fill."""
    from cppimpact_amd.surrogate import Surrogate, collect_samples

    x, conn = cube(8, jitter=0.1, seed=12, gap=1e-6)
    mat = ci.Elastoplastic(E, RHO, NU, 1.0, H, SY)
    smp = collect_samples(ci.Mesh.from_arrays(x, conn), mat,
                          ci.Wall("Wall", 0.0, wall_k(0.25, 6), None, 0, 1), [30.0], 0.0, 60)
    net = Surrogate(mat, seed=1)
    net.train(smp, epochs=4, batch=512)
    out = []
    for fill in (0, 1):
        d = ci.Dynamics(ci.Mesh.from_arrays(x, conn), mat,
                        ci.Wall("Wall", 0.0, wall_k(0.25, 6), None, 0, 1),
                        physics=L.PHYSICS_EP, precision=L.PREC_F32,
                        stress_update=L.RETURN_MLP, block_elements=64, debug_fill_nan=fill)
        d.set_surrogate(net)
        d.vel[:] = (0.0, 0.0, -30.0)
        d.init(0.0)
        d.step(60)
        xs, vs, *_ = d.fetch()
        _, ep = d.element_state()
        out.append((xs, vs, ep, d.surrogate_stats(), d.energy()[0]))
        d.close()
    assert out[0][3][0] > 0
    for a, b in zip(out[0][:3], out[1][:3]):
        assert np.all(np.isfinite(b)) and np.array_equal(a, b)
    assert out[0][3] == out[1][3]
    assert np.array_equal(out[0][4], out[1][4])
    net.close()
