"""Shared problem set-ups for the parity tests (the oracle and the HIP path
get the same seeded meshes, materials and walls)."""
import numpy as np

import cppimpact_amd as ci
from oracle import mlp_oracle as M
from oracle import pyoracle as O

E, NU, RHO, SY, H = 68.9e9, 0.33, 2700.0, 276e6, 700e6


def cube(n, quadratic=False, jitter=0.0, seed=0, gap=5e-5, side=0.25):
    x, conn = ci.box_mesh(n, n, n, side, side, side, origin=(0.0, 0.0, gap),
                          quadratic=quadratic, jitter=jitter, seed=seed)
    return x, conn


def wall_k(side, n):
    return 10.0 * E * side / n


def run_gpu(x, conn, v0, dt, nsteps, wall_stiffness=None, wall_location=0.0,
            slave_nodes=None, fixed_nodes=None, **opts):
    mesh = ci.Mesh.from_arrays(x, conn, slave_nodes=slave_nodes, fixed_nodes=fixed_nodes)
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


def plastic_states(rng, n, sy=SY, h=H):
    """Random stress states near the yield surface and strain increments
    that push most of them through it: (de, sigma_n, ep_n)."""
    ep = rng.uniform(0.0, 0.05, n)
    d = rng.normal(size=(n, 6))
    d[:, :3] -= d[:, :3].mean(axis=1, keepdims=True)
    d /= M.dev_norm(d)[:, None]
    radius = M.K23 * (sy + h * ep) * rng.uniform(0.9, 1.0, n)
    so = d * radius[:, None]
    so[:, :3] += rng.normal(0, 0.3 * sy, n)[:, None]
    de = rng.normal(0, 2e-4, (n, 6)) + 3e-4 * d
    return de, so, ep


def exact_samples(rng, n):
    """Samples (de, sigma_n, ep_n, sigma_{n+1}, ep_{n+1}) of the exact
    return at plastic_states, float32 like the recorder's."""
    de, so, ep = plastic_states(rng, n)
    tr = M.trial(de, so, E, NU)
    s1, ep1, _ = M.radial_return(tr, ep, E, NU, SY, H)
    keep = M.overshoot(tr, ep, SY, H) > 0
    smp = np.concatenate([de, so, ep[:, None], s1, ep1[:, None]], axis=1)[keep]
    return smp.astype(np.float32)
