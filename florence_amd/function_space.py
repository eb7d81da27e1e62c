"""FunctionSpace: basis tables at the quadrature points (host set-up).

Same constructor and attributes as the reference's FunctionSpace
(Florence/FunctionSpace/FunctionSpace.py:9-168) for the volume function space of
tetrahedra and hexahedra: ``Bases`` (nodes, G), ``Jm`` (3, nodes, G) -- the parent
gradients the C kernel indexes as ``Jm[d*G*npe + a*G + g]``
(_KinematicMeasures_.h:82-86) -- and ``AllGauss`` (G, 1).
"""
import numpy as np

from .bases import HexBases, TetBases

__all__ = ["FunctionSpace"]


class FunctionSpace(object):

    def __init__(self, mesh, quadrature=None, p=1, bases_type="nodal", bases_kind="CG",
                 evaluate_at_nodes=False, equally_spaced=False):
        if quadrature is None:
            raise ValueError("Function space requires a quadrature rule")
        if equally_spaced:
            raise NotImplementedError("equally spaced bases are outside the displacement-formulation path")
        self.ndim = 3
        self.element_type = mesh.element_type
        self.degree = mesh.InferPolynomialDegree()
        C = p - 1
        if self.degree - 1 != C:
            raise ValueError("Function space of the polynomial does not match element type")
        z = np.asarray(quadrature.points, dtype=np.float64)
        w = np.asarray(quadrature.weights, dtype=np.float64).ravel()
        if z.ndim == 1:
            raise ValueError("tensor-product quadrature must be flattened before building a function space")
        if mesh.element_type == "hex":
            N, dN = HexBases(C, z)
        elif mesh.element_type == "tet":
            N, dN = TetBases(C, z)
        else:
            raise NotImplementedError("function spaces are provided for tet and hex elements only")
        self.Bases = N
        self.gBasesx, self.gBasesy, self.gBasesz = dN[0], dN[1], dN[2]
        self.Jm = np.ascontiguousarray(dN)
        self.AllGauss = w[:, None].copy()
        self.quadrature = quadrature
        self.p = p

    @property
    def ngauss(self):
        return self.AllGauss.shape[0]

    @property
    def nodeperelem(self):
        return self.Bases.shape[0]
