"""Mesh / material builders shared by the tests (workloads of the assembly path)."""
import numpy as np

from florence_amd.function_space import FunctionSpace
from florence_amd.materials import MooneyRivlin, NeoHookean
from florence_amd.mesh import Mesh
from florence_amd.quadrature import QuadratureRule


def cube_mesh(element_type, p, n, jitter=0.0, seed=0):
    m = Mesh()
    m.Cube(n=n, element_type=element_type)
    m.GetHighOrderMesh(p=p)
    if jitter:
        m.JitterInteriorNodes(jitter / n, seed=seed, lower=(0, 0, 0), upper=(1, 1, 1))
    return m


def function_space(mesh, p, qtype="gauss", norder=None):
    if norder is None:
        norder = 2 * (p - 1) if mesh.element_type == "tet" else p + 1
        norder = max(norder, 1)
    q = QuadratureRule(qtype=qtype, norder=norder, mesh_type=mesh.element_type)
    return FunctionSpace(mesh, q, p=p)


def config1(n=8, seed=0):
    """Unit cube, n^3 cubes x 6 tets, p=2, degree-4 (11 point) rule, jitter +-0.05h, NH(1, 10)."""
    m = cube_mesh("tet", 2, n, jitter=0.05, seed=seed)
    fs = function_space(m, 2, norder=3)
    return m, fs, NeoHookean(3, mu=1.0, lamb=10.0)


def config2(n=24, seed=0):
    """Unit cube, n^3 hexes, p=4, 6x6x6 Gauss-Lobatto rule, jitter +-0.03h, MR(0.5, 0.5, 20)."""
    m = cube_mesh("hex", 4, n, jitter=0.03, seed=seed)
    fs = function_space(m, 4, qtype="lobatto", norder=6)
    return m, fs, MooneyRivlin(3, mu1=0.5, mu2=0.5, lamb=20.0)


def deformed(points, amp=0.05, seed=1):
    """A smooth finite deformation of the reference coordinates (plus small noise)."""
    X = np.asarray(points)
    rng = np.random.default_rng(seed)
    A = np.eye(3) + amp * rng.standard_normal((3, 3))
    x = X @ A.T + amp * np.sin(np.pi * X[:, [1, 2, 0]])
    return x
