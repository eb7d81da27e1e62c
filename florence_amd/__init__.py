"""florence_amd: the low-level displacement-formulation assembler of Florence and the
incremental Newton loop around it, on one MI355X.

The names a Florence user imports for this path (Florence/__init__.py re-exports them
from FiniteElements/Assembly/Assembly.py:25-95, Solver/FEMSolver.py,
Solver/LinearSolver.py, VariationalPrinciple/DisplacementFormulation.py,
MaterialLibrary/, MeshGeneration/Mesh.py, BoundaryCondition/BoundaryCondition.py,
QuadratureRules/QuadratureRule.py, FunctionSpace/FunctionSpace.py) are available at
package level.  Importing the package does not touch the GPU; the compute entry points
load libflorence_amd.so and raise if it or a HIP device is missing.
"""
from .assembly import Assemble, AssembleResidual, LowLevelAssembly
from .boundary_condition import BoundaryCondition
from .formulation import DisplacementFormulation
from .function_space import FunctionSpace
from .materials import (AnisotropicMooneyRivlin_1, LinearElastic, MooneyRivlin, NearlyIncompressibleMooneyRivlin,
                        NeoHookean)
from .mesh import Mesh
from .quadrature import QuadratureRule
from .solver import FEMSolver, LinearSolver
from .sparse import DeviceCSR

__all__ = ["Assemble", "AssembleResidual", "LowLevelAssembly", "FEMSolver", "LinearSolver",
           "DisplacementFormulation", "FunctionSpace", "QuadratureRule", "Mesh", "BoundaryCondition", "NeoHookean",
           "MooneyRivlin", "LinearElastic", "NearlyIncompressibleMooneyRivlin", "AnisotropicMooneyRivlin_1",
           "DeviceCSR"]
