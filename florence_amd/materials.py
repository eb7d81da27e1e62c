"""Hyperelastic materials with a low-level (device) dispatcher.

Constructor surface of the reference materials (kwargs become attributes,
Florence/MaterialLibrary/MaterialBase.py:11-70):

* ``NeoHookean(ndim, mu=, lamb=)`` -- Florence/MaterialLibrary/NeoHookean.py:5-32;
  device law restated from LLDispatch/CythonSource/_NeoHookean_.h:27-51.
* ``MooneyRivlin(ndim, mu1=, mu2=, lamb=)`` -- MooneyRivlin.py:8-34; device law
  restated from _MooneyRivlin_.h:30-70.

The stress/tangent formulas themselves live in the HIP kernels; these classes only
carry parameters and the identifiers the dispatcher uses
(FiniteElements/Assembly/_LowLevelAssembly_.py:47-60 dispatches on the class name).
"""
import numpy as np

__all__ = ["Material", "NeoHookean", "MooneyRivlin", "MATERIAL_IDS"]

# material identifiers of the C-ABI (include/florence_amd.h FL_MAT_*)
MATERIAL_IDS = {"NeoHookean": 1, "MooneyRivlin": 2}


class Material(object):

    def __init__(self, mtype, ndim, lame_parameter_1=None, lame_parameter_2=None, density=None, **kwargs):
        self.mtype = mtype
        self.ndim = ndim
        self.nvar = ndim
        self.mu = lame_parameter_1
        self.lamb = lame_parameter_2
        self.rho = density
        self.mu1 = None
        self.mu2 = None
        for k, v in kwargs.items():
            setattr(self, k, v)
        self.H_VoigtSize = 6 if ndim == 3 else 3
        self.fields = "mechanics"
        self.nature = "nonlinear"
        self.energy_type = "internal_energy"
        self.is_transversely_isotropic = False
        self.has_low_level_dispatcher = True

    @property
    def material_id(self):
        return MATERIAL_IDS[type(self).__name__]

    def device_parameters(self):
        """Packed parameter vector of the C-ABI: [mu, mu1, mu2, lamb]."""
        f = lambda v: 0.0 if v is None else float(v)
        return np.array([f(self.mu), f(self.mu1), f(self.mu2), f(self.lamb)], dtype=np.float64)


class NeoHookean(Material):
    """W(C) = mu/2 (C:I - 3) - mu ln J + lamb/2 (J-1)^2 (NeoHookean.py:5-10)."""

    def __init__(self, ndim, **kwargs):
        super(NeoHookean, self).__init__(type(self).__name__, ndim, **kwargs)
        if self.mu is None or self.lamb is None:
            raise ValueError("NeoHookean needs mu and lamb")


class MooneyRivlin(Material):
    """W(C) = mu1 C:I + mu2 G:I - 2(mu1+2 mu2) ln J + lamb/2 (J-1)^2 (MooneyRivlin.py:8-13)."""

    def __init__(self, ndim, **kwargs):
        super(MooneyRivlin, self).__init__(type(self).__name__, ndim, **kwargs)
        if self.mu1 is None or self.mu2 is None or self.lamb is None:
            raise ValueError("MooneyRivlin needs mu1, mu2 and lamb")
