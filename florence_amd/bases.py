"""Nodal interpolation bases on the reference hexahedron and tetrahedron.

Set-up code (host): evaluates N_a and dN_a/dxi at the quadrature points once per
(element type, p, rule); the device kernels only ever see the resulting tables.

Hexahedra -- tensor-product Lagrange bases through the Gauss-Lobatto-Legendre points
(Florence/FunctionSpace/ThreeDimensional/Hex/HexLagrangeGaussLobatto.py:5-79): the
tensor index is ``beta*n^2 + eta*n + zeta`` (first reference coordinate zeta fastest)
and nodes are numbered by ``NodeArrangementHex(C)[2]``
(Florence/MeshGeneration/NodeArrangement.py:201-211).

Tetrahedra -- the nodal basis of the complete polynomial space P_p through the
Fekete points of the reference tetrahedron
(Florence/FunctionSpace/ThreeDimensional/Tet/hpNodal.py:8-59, nodes from
FeketePointsTet.py).  The reference builds it from a Vandermonde matrix of
orthonormal Jacobi polynomials; the interpolant through a unisolvent node set is
unique, so any basis of P_p gives the same functions.  Here the spanning set is
products of Legendre polynomials P_i(x)P_j(y)P_k(z), i+j+k <= p, which keeps the
Vandermonde matrix well conditioned at p <= 4.  Gradients below 1e-14 in magnitude
are set to zero exactly as the reference does (GetBases.py:237 ``makezero``).
"""
from functools import lru_cache

import numpy as np
from numpy.polynomial import legendre as _leg

from ._tet_tables import FEKETE_TET
from .quadrature import GaussLobattoQuadrature

__all__ = [
    "LagrangeGaussLobatto1D", "NodeArrangementHex", "GaussLobattoPointsHex",
    "HexBases", "FeketePointsTet", "TetBases", "LinearTetBases", "LinearHexBases",
]


def LagrangeGaussLobatto1D(C, xi):
    """1D Lagrange bases through the C+2 GLL points and their derivatives at ``xi``
    (Florence/FunctionSpace/OneDimensional/Line.py:108-132).  Returns (N, dN), each (C+2,)."""
    n = C + 2
    nodes = GaussLobattoQuadrature(n)[0][:, 0]
    N = np.ones(n)
    dN = np.zeros(n)
    for a in range(n):
        others = [m for m in range(n) if m != a]
        denom = np.prod([nodes[a] - nodes[m] for m in others])
        N[a] = np.prod([xi - nodes[m] for m in others]) / denom
        s = 0.0
        for skip in others:
            s += np.prod([xi - nodes[m] for m in others if m != skip])
        dN[a] = s / denom
    return N, dN


@lru_cache(maxsize=None)
def NodeArrangementHex(C):
    """Local node -> tensor index map for a (C+2)^3-node hexahedron.

    Ordering (NodeArrangement.py:201-211): the four corners of the bottom layer,
    the four corners of the top layer, then layer by layer (beta ascending): the
    remaining nodes of the bottom layer, the four corners and then the rest of each
    intermediate layer, and the remaining nodes of the top layer.  Within a layer the
    corners are visited counter-clockwise (0, C+1, (C+2)^2-1, (C+2)^2-(C+2)) and the
    rest in increasing in-layer index ``eta*(C+2)+zeta``.
    """
    n = C + 2
    corners = [0, n - 1, n * n - 1, n * n - n]
    rest = [q for q in range(n * n) if q not in corners]
    layer = np.array(corners + rest, dtype=np.int64)
    order = [layer[:4], layer[:4] + (n - 1) * n * n, layer[4:]]
    for z in range(1, n - 1):
        order.append(layer + z * n * n)
    order.append(layer[4:] + (n - 1) * n * n)
    return np.concatenate(order).astype(np.int64)


def GaussLobattoPointsHex(C):
    """Reference-element coordinates (zeta, eta, beta) of the hex nodes in local
    order (Florence/QuadratureRules/GaussLobattoPoints.py:18-35)."""
    n = C + 2
    xs = GaussLobattoQuadrature(n)[0][:, 0]
    t = np.arange(n ** 3)
    beta, eta, zeta = t // (n * n), (t // n) % n, t % n
    pts = np.stack([xs[zeta], xs[eta], xs[beta]], axis=1)
    return pts[NodeArrangementHex(C)]


def HexBases(C, points):
    """Hex bases and parent gradients at ``points`` (npts, 3) = (zeta, eta, beta).

    Returns (N, dN) with N: (nodes, npts), dN: (3, nodes, npts) -- gradient with
    respect to (zeta, eta, beta) (HexLagrangeGaussLobatto.py:40-79).
    """
    n = C + 2
    perm = NodeArrangementHex(C)
    t = perm
    beta, eta, zeta = t // (n * n), (t // n) % n, t % n
    npts = points.shape[0]
    N = np.zeros((n ** 3, npts))
    dN = np.zeros((3, n ** 3, npts))
    for g in range(npts):
        Nz, dNz = LagrangeGaussLobatto1D(C, points[g, 0])
        Ne, dNe = LagrangeGaussLobatto1D(C, points[g, 1])
        Nb, dNb = LagrangeGaussLobatto1D(C, points[g, 2])
        N[:, g] = Nb[beta] * Ne[eta] * Nz[zeta]
        dN[0, :, g] = Nb[beta] * Ne[eta] * dNz[zeta]
        dN[1, :, g] = Nb[beta] * dNe[eta] * Nz[zeta]
        dN[2, :, g] = dNb[beta] * Ne[eta] * Nz[zeta]
    return N, dN


def LinearHexBases(points):
    """Trilinear hex bases (8, npts) in the p=1 node order of NodeArrangementHex(0)."""
    return HexBases(0, points)[0]


def FeketePointsTet(C):
    """Fekete nodes of the reference tetrahedron for p = C+1 (FeketePointsTet.py)."""
    if C not in FEKETE_TET:
        raise NotImplementedError("tetrahedral Fekete points are tabulated for p <= 4")
    return FEKETE_TET[C].copy()


def _exponents(p):
    return [(i, j, k) for s in range(p + 1) for i in range(s + 1) for j in range(s + 1 - i) for k in [s - i - j]]


def _legendre_table(p, x):
    """Legendre P_0..P_p and derivatives at x (vector)."""
    P = np.zeros((p + 1, x.shape[0]))
    dP = np.zeros((p + 1, x.shape[0]))
    for d in range(p + 1):
        c = [0.0] * d + [1.0]
        P[d] = _leg.legval(x, c)
        dP[d] = _leg.legval(x, _leg.legder(c))
    return P, dP


def _tet_span(p, pts):
    """Spanning set of P_p (Legendre products) and gradients at pts (npts, 3)."""
    ex = _exponents(p)
    Px, dPx = _legendre_table(p, pts[:, 0])
    Py, dPy = _legendre_table(p, pts[:, 1])
    Pz, dPz = _legendre_table(p, pts[:, 2])
    V = np.stack([Px[i] * Py[j] * Pz[k] for i, j, k in ex], axis=1)
    Vx = np.stack([dPx[i] * Py[j] * Pz[k] for i, j, k in ex], axis=1)
    Vy = np.stack([Px[i] * dPy[j] * Pz[k] for i, j, k in ex], axis=1)
    Vz = np.stack([Px[i] * Py[j] * dPz[k] for i, j, k in ex], axis=1)
    return V, Vx, Vy, Vz


def TetBases(C, points, makezero_tol=1e-14):
    """Nodal tet bases through FeketePointsTet(C) and parent gradients at ``points``.

    Returns (N, dN) with N: (nodes, npts), dN: (3, nodes, npts).
    """
    p = C + 1
    nodes = FeketePointsTet(C)
    Vn = _tet_span(p, nodes)[0]                    # (nodes, dim P_p)
    V, Vx, Vy, Vz = _tet_span(p, np.asarray(points, dtype=np.float64))
    # N_a(x) = sum_m coeff[m, a] phi_m(x) with Vn @ coeff = I
    rhs = np.concatenate([V, Vx, Vy, Vz], axis=0).T  # (dim, 4*npts)
    sol = np.linalg.solve(Vn.T, rhs)                 # (nodes, 4*npts)
    npts = V.shape[0]
    N = sol[:, :npts]
    dN = np.stack([sol[:, npts:2 * npts], sol[:, 2 * npts:3 * npts], sol[:, 3 * npts:]], axis=0)
    if makezero_tol:
        dN[np.abs(dN) < makezero_tol] = 0.0
    return N, dN


def LinearTetBases(points):
    """Linear tet bases (4, npts): vertices (-1,-1,-1), (1,-1,-1), (-1,1,-1), (-1,-1,1)."""
    x = np.asarray(points, dtype=np.float64)
    N1 = 0.5 * (1.0 + x[:, 0])
    N2 = 0.5 * (1.0 + x[:, 1])
    N3 = 0.5 * (1.0 + x[:, 2])
    return np.stack([1.0 - N1 - N2 - N3, N1, N2, N3], axis=0)
