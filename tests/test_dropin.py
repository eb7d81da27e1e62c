"""The drop-in Python surface, without a GPU: package-level names of the reference's
API for this path, and the refusals of Assembly.py:37-42 / _LowLevelAssembly_.py:45-60
(raised before any device work)."""
import numpy as np
import pytest

import florence_amd
from florence_amd import DisplacementFormulation, FEMSolver, Mesh, NeoHookean


def test_package_exports_reference_names():
    for name in ("Assemble", "LowLevelAssembly", "FEMSolver", "LinearSolver", "DisplacementFormulation",
                 "NeoHookean", "MooneyRivlin", "LinearElastic", "NearlyIncompressibleMooneyRivlin",
                 "AnisotropicMooneyRivlin_1", "Mesh", "BoundaryCondition", "QuadratureRule", "FunctionSpace"):
        assert hasattr(florence_amd, name), name
        assert name in florence_amd.__all__


def _tiny():
    m = Mesh()
    m.Cube(n=1, element_type="tet")
    m.GetHighOrderMesh(p=2)
    return m, DisplacementFormulation(m)


def test_material_without_dispatcher_is_refused():
    m, form = _tiny()
    mat = NeoHookean(3, mu=1.0, lamb=10.0)
    mat.has_low_level_dispatcher = False
    with pytest.raises(RuntimeError, match="low level"):
        florence_amd.Assemble(FEMSolver(optimise=True), form.function_spaces[0], form, m, mat,
                              np.asarray(m.points), None)


def test_unknown_material_is_refused():
    class IncrementalLinearElastic(NeoHookean):
        pass

    m, form = _tiny()
    mat = IncrementalLinearElastic(3, mu=1.0, lamb=10.0)
    with pytest.raises(NotImplementedError, match="IncrementalLinearElastic"):
        florence_amd.Assemble(FEMSolver(optimise=True), form.function_spaces[0], form, m, mat,
                              np.asarray(m.points), None)
