"""Synthetic workloads of the assembly path: seeded meshes, rules, materials and
boundary conditions of the benchmark workloads.

* ``cube_tet_p2``  -- unit cube, n^3 cubes split into 6 tets, p=2 (10 nodes), the
  degree-4 (11 point) tetrahedral rule, interior nodes jittered by U(-0.05h, 0.05h)
  with h = 1/n, the linear cell edge; Neo-Hookean(mu=1, lamb=10); x=0 clamped, x=1
  given u_x=0.2 (5 load increments).
* ``cube_hex_p4``  -- unit cube, n^3 hexes, p=4 (125 nodes), 6x6x6 Gauss-Lobatto rule,
  interior nodes jittered by U(-0.03h, 0.03h) with h = 1/(n p), the node spacing along
  an element edge; Mooney-Rivlin(mu1=0.5, mu2=0.5, lamb=20); x=0 clamped, x=1 twisted
  30 degrees about the x axis through (y, z) = (0.5, 0.5) (10 load increments).

"Interior" means strictly inside the cube so the boundary faces stay planar.  For the
p=4 hexes h is the node spacing: taking h as the cell edge 1/n folds the elements
(+-0.03/n moves the Gauss-Lobatto nodes next to a corner enough to make det(dX/dxi)
negative at 1,479 corner points of the 24^3 mesh).  The p=2 tets keep h = 1/n: their
mid-edge nodes sit 1/(2n) apart, ten times the jitter.  All randomness is numpy
PCG64(seed).
"""
import numpy as np

from .function_space import FunctionSpace
from .materials import MooneyRivlin, NeoHookean
from .mesh import Mesh
from .quadrature import QuadratureRule

__all__ = ["cube_mesh", "cube_tet_p2", "cube_hex_p4", "clamp_and_pull", "clamp_and_twist", "twist_field"]


def cube_mesh(element_type, p, n, jitter=0.0, seed=0, spacing="node", numbering="first"):
    """Unit cube with n^3 linear cells (hexes, or 6 tets each), raised to degree p
    (node numbering: "first" appearance, or the reference's, see Mesh.GetHighOrderMesh);
    ``jitter`` is the half-width of the uniform noise in units of h: the node spacing
    1/(n p) (``spacing="node"``) or the cell edge 1/n (``spacing="cell"``), drawn per node
    in node order."""
    m = Mesh()
    m.Cube(n=n, element_type=element_type)
    m.GetHighOrderMesh(p=p, numbering=numbering)
    if jitter:
        h = 1.0 / n if spacing == "cell" else 1.0 / (n * p)
        m.JitterInteriorNodes(jitter * h, seed=seed, lower=(0, 0, 0), upper=(1, 1, 1))
    return m


def cube_tet_p2(n=8, seed=0):
    m = cube_mesh("tet", 2, n, jitter=0.05, seed=seed, spacing="cell")
    fs = FunctionSpace(m, QuadratureRule(norder=3, mesh_type="tet"), p=2)
    return m, fs, NeoHookean(3, mu=1.0, lamb=10.0)


def cube_hex_p4(n=24, seed=0, numbering="first"):
    m = cube_mesh("hex", 4, n, jitter=0.03, seed=seed, numbering=numbering)
    fs = FunctionSpace(m, QuadratureRule(qtype="lobatto", norder=6, mesh_type="hex"), p=4)
    return m, fs, MooneyRivlin(3, mu1=0.5, mu2=0.5, lamb=20.0)


def clamp_and_pull(mesh, ux=0.2, tol=1e-10):
    """Dirichlet flags (nnode, 3), NaN = free: x=0 fully clamped, x=1 u_x = ux."""
    X = np.asarray(mesh.points)
    flags = np.full(X.shape, np.nan)
    flags[np.abs(X[:, 0]) < tol, :] = 0.0
    flags[np.abs(X[:, 0] - 1.0) < tol, 0] = ux
    return flags


def twist_field(X, theta, center=(0.5, 0.5)):
    """Displacement of a rigid rotation by theta about the x axis through (y,z)=center."""
    c, s = np.cos(theta), np.sin(theta)
    y, z = X[:, 1] - center[0], X[:, 2] - center[1]
    u = np.zeros_like(X)
    u[:, 1] = c * y - s * z - y
    u[:, 2] = s * y + c * z - z
    return u


def clamp_and_twist(mesh, degrees=30.0, tol=1e-10):
    """Dirichlet flags (nnode, 3): x=0 clamped, x=1 face rotated about the x axis."""
    X = np.asarray(mesh.points)
    flags = np.full(X.shape, np.nan)
    flags[np.abs(X[:, 0]) < tol, :] = 0.0
    right = np.abs(X[:, 0] - 1.0) < tol
    flags[right, :] = twist_field(X[right], np.deg2rad(degrees))
    return flags


def twisted_state(mesh, degrees=15.0):
    """A smooth synthetic current configuration: the x=1 twist interpolated linearly
    in x (plus its natural x-shortening left out).  Used as the deformed state the
    assembly benchmark evaluates K and T at."""
    X = np.asarray(mesh.points)
    u = twist_field(X, np.deg2rad(degrees)) * X[:, [0]]
    return X + u
