"""Hyperelastic materials with a low-level (device) dispatcher.

Constructor surface of the reference materials (kwargs become attributes,
Florence/MaterialLibrary/MaterialBase.py:11-70):

* ``NeoHookean(ndim, mu=, lamb=)`` -- Florence/MaterialLibrary/NeoHookean.py:5-32;
  device law restated from LLDispatch/CythonSource/_NeoHookean_.h:27-51.
* ``MooneyRivlin(ndim, mu1=, mu2=, lamb=)`` -- MooneyRivlin.py:8-34; device law
  restated from _MooneyRivlin_.h:30-70.
* ``LinearElastic(ndim, mu=, lamb=)`` -- LinearElastic.py:6-26; device law restated
  from _LinearElastic_.h:26-52, assembled without the geometric stiffness as the
  generated low-level assembler is (AOT_Assembler.py:84-93).
* ``NearlyIncompressibleMooneyRivlin(ndim, mu=, lamb=, mu1=, mu2=)`` --
  NearlyIncompressibleMooneyRivlin.py:14-43; device law restated from
  _NearlyIncompressibleMooneyRivlin_.h:28-87.  The Python class derives alpha, beta,
  kappa from mu and lamb (:30-33), but the low-level assembler is built on
  (mu1, mu2, lamb) as (alpha, beta, kappa) (AOT_Assembler.py:69-70, 114-115); the
  device path follows the low-level assembler, so mu1 and mu2 must be given.
* ``AnisotropicMooneyRivlin_1(ndim, mu1=, mu2=, mu3=, lamb=,
  anisotropic_orientations=)`` -- AnisotropicMooneyRivlin_1.py:8-37; device law
  restated from _AnisotropicMooneyRivlin_1_.h:30-95 with the element's fibre
  direction anisotropic_orientations[elem] (AOT_Assembler.py:78-82).

The stress/tangent formulas themselves live in the HIP kernels; these classes only
carry parameters and the identifiers the dispatcher uses
(FiniteElements/Assembly/_LowLevelAssembly_.py:47-60 dispatches on the class name).
"""
import numpy as np

__all__ = ["Material", "NeoHookean", "MooneyRivlin", "LinearElastic", "NearlyIncompressibleMooneyRivlin",
           "AnisotropicMooneyRivlin_1", "MATERIAL_IDS"]

# material identifiers of the C-ABI (include/florence_amd.h FL_MAT_*)
MATERIAL_IDS = {"NeoHookean": 1, "MooneyRivlin": 2, "LinearElastic": 3, "NearlyIncompressibleMooneyRivlin": 4,
                "AnisotropicMooneyRivlin_1": 5}


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
        self.mu3 = None
        self.anisotropic_orientations = None
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
        """Packed parameter vector of the C-ABI: [mu, mu1, mu2, lamb, mu3]."""
        f = lambda v: 0.0 if v is None else float(v)
        return np.array([f(self.mu), f(self.mu1), f(self.mu2), f(self.lamb), f(self.mu3)], dtype=np.float64)

    def device_table(self, nelem):
        """(elem_material int32[nelem], table f64[nelem][stride]) for laws with
        per-element data, or None: element e uses row e, {mu, mu1, mu2, lamb, mu3, N}."""
        return None

    def device_table_row(self):
        """This law's row of a per-element material table of the C-ABI (several
        materials of one law in one launch): {mu, mu1, mu2, lamb}, the first four
        entries of device_parameters (fl_material_table_stride = 4)."""
        return self.device_parameters()[:4]


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


class LinearElastic(Material):
    """W = lamb/2 tr(e)^2 + mu e:e (LinearElastic.py:6-11)."""

    def __init__(self, ndim, **kwargs):
        super(LinearElastic, self).__init__(type(self).__name__, ndim, **kwargs)
        if self.mu is None or self.lamb is None:
            raise ValueError("LinearElastic needs mu and lamb")
        self.nature = "linear"


class NearlyIncompressibleMooneyRivlin(Material):
    """W = alpha J^(-2/3) C:I + beta J^-2 (G:I)^(3/2) + kappa/2 (J-1)^2
    (NearlyIncompressibleMooneyRivlin.py:14-23)."""

    def __init__(self, ndim, **kwargs):
        super(NearlyIncompressibleMooneyRivlin, self).__init__(type(self).__name__, ndim, **kwargs)
        if self.mu is not None and self.lamb is not None:
            # the Python model's own constants (NearlyIncompressibleMooneyRivlin.py:30-33)
            self.gamma = 1.
            self.alpha = self.gamma * self.mu / 2.
            self.beta = (self.mu - 2. * self.alpha) / 3. / np.sqrt(3.)
            self.kappa = self.lamb + 4.0 / 3.0 * self.alpha + 2.0 * np.sqrt(3.0) * self.beta
        if self.mu1 is None or self.mu2 is None or self.lamb is None:
            raise ValueError("the low-level NearlyIncompressibleMooneyRivlin takes mu1, mu2 and lamb as "
                             "alpha, beta and kappa (AOT_Assembler.py:69-70)")


class AnisotropicMooneyRivlin_1(Material):
    """W = mu1 C:I + mu2 G:I - 2(mu1+2mu2+mu3) ln J + lamb/2 (J-1)^2 + mu3 FN.FN + mu3 HN.HN
    (AnisotropicMooneyRivlin_1.py:8-15); one fibre direction N per element."""

    def __init__(self, ndim, **kwargs):
        super(AnisotropicMooneyRivlin_1, self).__init__(type(self).__name__, ndim, **kwargs)
        self.is_transversely_isotropic = True
        if self.mu1 is None or self.mu2 is None or self.mu3 is None or self.lamb is None:
            raise ValueError("AnisotropicMooneyRivlin_1 needs mu1, mu2, mu3 and lamb")

    def device_table(self, nelem):
        ori = self.anisotropic_orientations
        if ori is None:
            raise ValueError("AnisotropicMooneyRivlin_1 needs anisotropic_orientations (nelem, 3)")
        ori = np.asarray(ori, dtype=np.float64).reshape(-1, 3)
        if ori.shape[0] != nelem:
            raise ValueError("anisotropic_orientations has {} rows for {} elements".format(ori.shape[0], nelem))
        table = np.empty((nelem, 8), dtype=np.float64)
        table[:, :5] = self.device_parameters()
        table[:, 5:] = ori
        return np.arange(nelem, dtype=np.int32), table
