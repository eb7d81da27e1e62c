"""Quadrature rules for the displacement-formulation assembler set-up.

Host-side tables only: these arrays are computed once per mesh and handed to the
device kernels, exactly as the reference hands ``AllGauss``/``Jm`` to its C
kernel.

* Gauss-Legendre rules come from ``numpy.polynomial.legendre.leggauss`` -- the same
  library routine the reference calls (Florence/QuadratureRules/
  NumericIntegrator.py:7-9).
* Gauss-Lobatto-Legendre rules are computed from their closed form (interior nodes
  are the roots of P'_{n-1}; weights 2 / (n (n-1) P_{n-1}(x)^2)); the reference
  tabulates the same numbers (NumericIntegrator.py:93-135).
* Tetrahedral rules are the reference's published tables (QuadraturePointsWeightsTet
  with Opt=3, QuadraturePointsWeightsTet.py:32-160).
* Tensor-product rules are flattened in i-j-k order, first coordinate slowest
  (QuadratureRule.py:99-113).
"""
import numpy as np
from numpy.polynomial import legendre as _leg

from ._tet_tables import TET_QUADRATURE

__all__ = ["GaussQuadrature", "GaussLobattoQuadrature", "QuadratureRule"]


def GaussQuadrature(N, a=-1.0, b=1.0):
    """N-point Gauss-Legendre points and weights on [a, b] (NumericIntegrator.py:7-9)."""
    z, w = _leg.leggauss(int(N))
    if a == -1 and b == 1:
        return z, w
    return 0.5 * (b - a) * z + 0.5 * (b + a), 0.5 * (b - a) * w


def _legendre_and_derivative(n, x):
    """P_n(x) and P'_n(x) by the three-term recurrence (x inside (-1, 1))."""
    p0, p1 = np.ones_like(x), x.copy()
    if n == 0:
        return p0, np.zeros_like(x)
    for k in range(1, n):
        p0, p1 = p1, ((2 * k + 1) * x * p1 - k * p0) / (k + 1)
    dp = n * (x * p1 - p0) / (x * x - 1.0)
    return p1, dp


def GaussLobattoQuadrature(n):
    """n-point Gauss-Lobatto-Legendre rule on [-1, 1].

    Returns ``(z, w)`` as column arrays of shape (n, 1), the layout of the reference's
    ``GaussLobattoQuadrature(C)`` (NumericIntegrator.py:93-135, called with the
    number of points).  Interior nodes are the roots of P'_{n-1}, found by Newton's
    method from Chebyshev-Gauss-Lobatto guesses; weights are 2/(n(n-1)P_{n-1}^2).
    """
    n = int(n)
    if n < 2:
        raise ValueError("There should be at least two Gauss-Lobatto points")
    N = n - 1
    x = -np.cos(np.pi * np.arange(n) / N)
    interior = x[1:-1].copy()
    for _ in range(100):
        # Newton on q(x) = P'_N(x); q' from the Legendre ODE: (1-x^2) P''_N = 2x P'_N - N(N+1) P_N
        p, dp = _legendre_and_derivative(N, interior)
        d2p = (2.0 * interior * dp - N * (N + 1) * p) / (1.0 - interior * interior)
        step = dp / d2p
        interior -= step
        if np.max(np.abs(step), initial=0.0) < 1e-16:
            break
    x[1:-1] = interior
    x[0], x[-1] = -1.0, 1.0
    # symmetrise (the exact rule is symmetric about 0)
    x = 0.5 * (x - x[::-1])
    if n % 2 == 1:
        x[n // 2] = 0.0
    pN = np.empty_like(x)
    for i, xi in enumerate(x):
        pN[i] = _leg.legval(xi, [0] * N + [1])
    w = 2.0 / (N * (N + 1) * pN * pN)
    return x[:, None], w[:, None]


class QuadratureRule(object):
    """Points and weights on the reference element, same constructor as the
    reference (QuadratureRule.py:12-78).

    ``mesh_type`` "hex": Gauss-Legendre with ``norder`` points per direction,
    flattened (QuadratureRule.py:41-50, 99-113).  ``qtype="lobatto"`` gives the
    Gauss-Lobatto-Legendre tensor rule instead (the reference reaches it through a
    user supplied ``quadrature_rules``).
    ``mesh_type`` "tet": the tabulated rule of order ``norder``
    (QuadraturePointsWeightsTet.py:32-160).
    """

    def __init__(self, qtype="gauss", norder=2, mesh_type="tri", optimal=3, flatten=True, evaluate=True):
        self.qtype = qtype
        self.norder = int(norder)
        self.element_type = mesh_type
        self.optimal = optimal
        self.flatten = flatten
        self.points, self.weights = [], []
        if not evaluate:
            return
        if mesh_type == "hex":
            if qtype == "lobatto":
                z, w = GaussLobattoQuadrature(self.norder)
                z, w = z[:, 0], w[:, 0]
            else:
                z, w = GaussQuadrature(self.norder)
            self.points, self.weights = z, w
            if flatten:
                self.Flatten(mesh_type)
        elif mesh_type == "tet":
            key = max(self.norder, 1)
            if key not in TET_QUADRATURE:
                raise NotImplementedError("tetrahedral quadrature of order {} is not tabulated".format(norder))
            zw = TET_QUADRATURE[key]
            self.points = zw[:, :3].copy()
            self.weights = zw[:, 3].copy()
        elif mesh_type == "line":
            self.points, self.weights = GaussQuadrature(self.norder)
        else:
            raise NotImplementedError("element type {} is outside the displacement-formulation path".format(mesh_type))

    def Flatten(self, mesh_type=None):
        """Tensor rule -> flattened (npts**3, 3) points, i-j-k order (QuadratureRule.py:99-113)."""
        z, w = np.asarray(self.points), np.asarray(self.weights)
        n = z.shape[0]
        ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
        ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
        self.weights = w[ii] * w[jj] * w[kk]
        self.points = np.stack([z[ii], z[jj], z[kk]], axis=1)
