"""DisplacementFormulation: quadrature + function space set-up of the DF path.

Constructor surface of Florence/VariationalPrinciple/DisplacementFormulation.py:14-38
with the quadrature-order rule of VariationalPrinciple.py:53-72 (tets: 2C, hexes:
C+2 Gauss points per direction, or ``quadrature_degree``) and the
``quadrature_rules`` override (VariationalPrinciple.py:85-107).
"""
from .function_space import FunctionSpace
from .quadrature import QuadratureRule

__all__ = ["DisplacementFormulation"]


class DisplacementFormulation(object):

    def __init__(self, mesh, variables_order=(1,), quadrature_rules=None, quadrature_type=None,
                 function_spaces=None, compute_post_quadrature=False, equally_spaced_bases=False,
                 quadrature_degree=None):
        if mesh.element_type not in ("tet", "hex"):
            raise NotImplementedError("DisplacementFormulation is provided for tet and hex elements")
        self.ndim = 3
        self.nvar = 3
        self.fields = "mechanics"
        self.variables_order = variables_order
        C = mesh.InferPolynomialDegree() - 1
        if quadrature_rules is None:
            norder = self.GetQuadratureOrder(C, mesh.element_type, quadrature_degree)
            quadrature = QuadratureRule(optimal=3, norder=norder, mesh_type=mesh.element_type)
            self.quadrature_rules = (quadrature, None, None)
        else:
            self.quadrature_rules = quadrature_rules if isinstance(quadrature_rules, (tuple, list)) \
                else (quadrature_rules, None, None)
        if function_spaces is None:
            fs = FunctionSpace(mesh, self.quadrature_rules[0], p=C + 1, equally_spaced=equally_spaced_bases)
            self.function_spaces = (fs, None, None)
        else:
            self.function_spaces = function_spaces

    @staticmethod
    def GetQuadratureOrder(C, element_type, quadrature_degree=None):
        if quadrature_degree is not None:
            return quadrature_degree
        if element_type in ("tet", "tri"):
            return 2 * C if C > 0 else 1
        return C + 2
