"""Shared problem set-ups for the parity tests (the oracle and the HIP path
get the same seeded meshes, materials and walls)."""
import numpy as np

import cppimpact_amd as ci
from oracle import pyoracle as O

E, NU, RHO, SY, H = 68.9e9, 0.33, 2700.0, 276e6, 700e6


def cube(n, quadratic=False, jitter=0.0, seed=0, gap=5e-5, side=0.25):
    x, conn = ci.box_mesh(n, n, n, side, side, side, origin=(0.0, 0.0, gap),
                          quadratic=quadratic, jitter=jitter, seed=seed)
    return x, conn


def wall_k(side, n):
    return 10.0 * E * side / n


def run_gpu(x, conn, v0, dt, nsteps, wall_stiffness=None, wall_location=0.0, **opts):
    mesh = ci.Mesh.from_arrays(x, conn)
    mat = ci.Elastoplastic(E, RHO, NU, 1.0, H, SY)
    wall = None
    if wall_stiffness is not None:
        wall = ci.Wall("Wall", wall_location, wall_stiffness, None, 0, 1)
    dyn = ci.Dynamics(mesh, mat, wall, **opts)
    dyn.vel[:] = np.asarray(v0, dtype=np.float64)
    dyn.init(dt)
    if nsteps:
        dyn.step(nsteps)
    out = dict(zip(("x", "v", "a", "m", "v_i"), dyn.fetch()))
    out["energy"], out["step"] = dyn.energy()
    out["dyn"] = dyn
    return out


def run_oracle(x, conn, v0, dt, nsteps, physics, wall_stiffness=None, wall_location=0.0,
               **kw):
    wall = None
    if wall_stiffness is not None:
        wall = dict(location=wall_location, stiffness=wall_stiffness)
    s = O.Solver(x, conn, v0, E=E, nu=NU, rho=RHO, sigma_y=SY, H=H, physics=physics,
                 wall=wall, dt=dt, **kw)
    return s.solve(nsteps)


def rel(a, b, scale=None):
    a, b = np.asarray(a), np.asarray(b)
    s = np.abs(b).max() if scale is None else scale
    return float(np.abs(a - b).max() / max(s, 1e-300))
