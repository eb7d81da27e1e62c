"""BoundaryCondition: nodal Dirichlet data of the static DF path.

Surface of Florence/BoundaryCondition/BoundaryCondition.py used by the Newton loop:
``SetDirichletCriteria(func, *args)`` (:233-246) stores a (nnode, 3) array with NaN
for free dofs; ``GetDirichletBoundaryConditions`` (:303-409, 'straight' nodal branch
:362-392) derives ``columns_out`` (flat constrained dofs), ``applied_dirichlet`` and
``columns_in``.  ``GetReducedMatrices`` / ``UpdateFixDoFs`` / ``UpdateFreeDoFs``
(:842-933) become a device mask of free dofs used by the masked operator of
libflorence_amd's solver, so K is never sliced.
"""
import numpy as np

__all__ = ["BoundaryCondition"]


class BoundaryCondition(object):

    def __init__(self):
        self.dirichlet_flags = None
        self.neumann_flags = None
        self.columns_in = None
        self.columns_out = None
        self.applied_dirichlet = None
        self.boundary_type = "straight"

    def SetDirichletCriteria(self, func, *args, **kwargs):
        self.dirichlet_flags = func(*args, **kwargs)
        return self.dirichlet_flags

    def GetDirichletBoundaryConditions(self, formulation, mesh, material=None, solver=None, fem_solver=None):
        if self.dirichlet_flags is None:
            raise RuntimeError("Dirichlet boundary conditions are not set for the analysis")
        flat = np.asarray(self.dirichlet_flags, dtype=np.float64).ravel()
        self.columns_out = np.arange(flat.size)[~np.isnan(flat)].astype(np.int64)
        self.applied_dirichlet = flat[~np.isnan(flat)]
        self.columns_in = np.delete(np.arange(formulation.nvar * mesh.points.shape[0]), self.columns_out)
        return self.columns_in, self.columns_out, self.applied_dirichlet

    def free_mask(self, ndof):
        m = np.ones(ndof, dtype=np.uint8)
        m[self.columns_out] = 0
        return m

    def ComputeNeumannForces(self, mesh, *args, **kwargs):
        """No surface or body loads on this path: zero nodal forces."""
        return np.zeros((3 * mesh.points.shape[0], 1))
