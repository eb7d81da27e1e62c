"""Mesh / material builders shared by the tests (workloads of the assembly path)."""
import numpy as np

from florence_amd.function_space import FunctionSpace
from florence_amd.quadrature import QuadratureRule
from florence_amd.workloads import cube_hex_p4, cube_mesh, cube_tet_p2  # noqa: F401


def function_space(mesh, p, qtype="gauss", norder=None):
    if norder is None:
        norder = 2 * (p - 1) if mesh.element_type == "tet" else p + 1
        norder = max(norder, 1)
    q = QuadratureRule(qtype=qtype, norder=norder, mesh_type=mesh.element_type)
    return FunctionSpace(mesh, q, p=p)


def tet_p2_cube(n=8, seed=0):
    """Unit cube, n^3 cubes x 6 tets, p=2, degree-4 (11 point) rule, jitter +-0.05h, NH(1, 10)."""
    return cube_tet_p2(n, seed)


def hex_p4_cube(n=24, seed=0):
    """Unit cube, n^3 hexes, p=4, 6x6x6 Gauss-Lobatto rule, jitter +-0.03h, MR(0.5, 0.5, 20)."""
    return cube_hex_p4(n, seed)


def deformed(points, amp=0.05, seed=1):
    """A smooth finite deformation of the reference coordinates."""
    X = np.asarray(points)
    rng = np.random.default_rng(seed)
    A = np.eye(3) + amp * rng.standard_normal((3, 3))
    return X @ A.T + amp * np.sin(np.pi * X[:, [1, 2, 0]])
