"""The benchmark configurations of the explicit-dynamics step.

One place for the meshes, walls, impact directions and run lengths that
bench.py, the GPU tests and the surrogate's training set use:

  c0  10^3-hex cube, 6,000 tet4, -10 m/s, 1,000 steps
  c1  20^3-hex cube, 48,000 tet10 (4-point rule), -20 m/s, 2,000 steps
  c2  64^3-hex cube, 1,572,864 tet4, interior nodes jittered by 0.1 of the
      spacing (SplitMix64 seed 0), -10 m/s, 500 steps
  c3  thin-walled cylinder R 1 m, L 4 m, t 0.02 m, 360 x 185 x 3 cells x 6
      tet10 (1,198,800 elements), oblique impact 30 degrees from the wall
      normal at 7 m/s, 1,000 steps
  c4  the c2 mesh with the MLP surrogate of the return map

All share the aluminium of the configurations (E 68.9 GPa, nu 0.33, rho
2700 kg/m^3, Y0 276 MPa, linear hardening H 700 MPa) and a rigid wall at
z = 0 with penalty stiffness k = 10 E h (h the cell size through the
wall-side thickness), as the reference's driver builds it (test.cu:57-95).
"""
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L
from .dynamics import Dynamics, Elastoplastic, Mesh, Wall

E, NU, RHO, SY, H = 68.9e9, 0.33, 2700.0, 276e6, 700e6
SIDE = 0.25


def material():
    return Elastoplastic(E, RHO, NU, 1.0, H, SY)


@dataclass
class Workload:
    name: str
    x: np.ndarray            # [N][3] reference positions
    conn: np.ndarray         # [E][npe]
    wall_stiffness: float
    direction: tuple         # unit impact direction of v0
    speed: float             # m/s
    steps: int
    options: dict = field(default_factory=dict)

    @property
    def elements(self):
        return int(self.conn.shape[0])

    def dynamics(self, speed=None, **options):
        """A Dynamics context of this workload (J2 physics), v0 set."""
        opts = dict(physics=L.PHYSICS_EP)
        opts.update(self.options)
        opts.update(options)
        d = Dynamics(Mesh.from_arrays(self.x, self.conn), material(),
                     Wall("Wall", 0.0, self.wall_stiffness, None, 0, 1), **opts)
        v = self.speed if speed is None else speed
        d.vel[:] = np.asarray(self.direction, dtype=np.float64) * float(v)
        return d


def c0():
    x, conn = L.box_mesh(10, 10, 10, SIDE, SIDE, SIDE, origin=(0.0, 0.0, 5e-5))
    return Workload("c0", x, conn, 10 * E * SIDE / 10, (0.0, 0.0, -1.0), 10.0, 1000)


def c1():
    x, conn = L.box_mesh(20, 20, 20, SIDE, SIDE, SIDE, origin=(0.0, 0.0, 1e-6), quadratic=True)
    return Workload("c1", x, conn, 10 * E * SIDE / 20, (0.0, 0.0, -1.0), 20.0, 2000)


def c2(n=64):
    x, conn = L.box_mesh(n, n, n, SIDE, SIDE, SIDE, origin=(0.0, 0.0, 1e-6), jitter=0.1, seed=0)
    return Workload("c2", x, conn, 10 * E * SIDE / n, (0.0, 0.0, -1.0), 10.0, 500)


def c3():
    x, conn = L.cylinder_mesh(360, 185, 3, 1.0, 4.0, 0.02, origin=(0.0, 0.0, 1e-6),
                              quadratic=True)
    a = math.radians(30.0)
    return Workload("c3", x, conn, 10 * E * 0.02 / 3, (math.sin(a), 0.0, -math.cos(a)), 7.0,
                    1000)
