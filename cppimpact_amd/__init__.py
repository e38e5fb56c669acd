"""MI355X (gfx950) explicit-dynamics step of cppimpact behind the reference's
Mesh / Elastoplastic / Wall / Dynamics interface."""
from . import _lib
from ._lib import CpiError, box_mesh, cylinder_mesh, greedy_coloring
from .dynamics import BaseMaterial, Dynamics, Elastoplastic, Mesh, Wall, loop_iterations

__all__ = ["CpiError", "box_mesh", "cylinder_mesh", "greedy_coloring", "BaseMaterial",
           "Dynamics", "Elastoplastic", "Mesh", "Wall", "loop_iterations", "_lib"]
