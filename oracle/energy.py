"""Independent energy oracle -- TEST INFRASTRUCTURE (see oracle/__init__.py).

Total potential energy of a small mesh in torch float64, with derivatives by
autograd:

    Pi(x) = sum_e sum_g w_g |det(dX/dxi)| W(F_g),   F = dx/dX

with the strain energies stated by the reference's material docstrings and
``InternalEnergy`` methods:
  Neo-Hookean  W = mu/2 (tr C - 3) - mu ln J + lamb/2 (J-1)^2        (NeoHookean.py:70-84)
  Mooney-Rivlin W = mu1 (tr C - 3) + mu2 (tr G - 3) - 2(mu1+2mu2) ln J + lamb/2 (J-1)^2,
                G = H^T H, H = J F^{-T}                              (MooneyRivlin.py:80-95)
  Nearly incompressible Mooney-Rivlin
                W = alpha J^(-2/3) tr C + beta J^-2 (tr G)^(3/2) + kappa/2 (J-1)^2
                (NearlyIncompressibleMooneyRivlin.py:17-19), with (alpha, beta, kappa)
                = (mu1, mu2, lamb) as the low-level assembler takes them
  Anisotropic Mooney-Rivlin 1
                W = mu1 tr C + mu2 tr G - 2(mu1+2mu2+mu3) ln J + lamb/2 (J-1)^2
                    + mu3 FN.FN + mu3 HN.HN              (AnisotropicMooneyRivlin_1.py:9-15)
  Linear elastic (small strain, reference configuration)
                W = mu e:e + lamb/2 tr(e)^2, e = sym(F) - I       (LinearElastic.py:8-10)

dPi/dx must equal the assembled internal force T and d2Pi/dx2 the assembled
tangent K (updated-Lagrangian integration is exact per quadrature point for
det(dx/dxi) > 0), which pins the kinematics, the material laws, the Voigt
convention, B, the geometric term and the scatter at once.
"""
import numpy as np
import torch


def strain_energy(F, material, N=None):
    name = type(material).__name__ if not isinstance(material, dict) else material["name"]
    p = material if isinstance(material, dict) else material.__dict__
    if name == "LinearElastic":
        e = 0.5 * (F + F.transpose(-1, -2)) - torch.eye(3, dtype=F.dtype)
        tre = e.diagonal(dim1=-2, dim2=-1).sum(-1)
        return p["mu"] * (e * e).sum((-1, -2)) + p["lamb"] / 2.0 * tre ** 2
    J = torch.linalg.det(F)
    C = F.transpose(-1, -2) @ F
    trC = C.diagonal(dim1=-2, dim2=-1).sum(-1)
    if name == "NeoHookean":
        mu, lamb = p["mu"], p["lamb"]
        return mu / 2.0 * (trC - 3.0) - mu * torch.log(J) + lamb / 2.0 * (J - 1.0) ** 2
    if name == "MooneyRivlin":
        mu1, mu2, lamb = p["mu1"], p["mu2"], p["lamb"]
        H = J[..., None, None] * torch.linalg.inv(F).transpose(-1, -2)
        G = H.transpose(-1, -2) @ H
        trG = G.diagonal(dim1=-2, dim2=-1).sum(-1)
        return mu1 * (trC - 3.0) + mu2 * (trG - 3.0) - 2.0 * (mu1 + 2.0 * mu2) * torch.log(J) \
            + lamb / 2.0 * (J - 1.0) ** 2
    H = J[..., None, None] * torch.linalg.inv(F).transpose(-1, -2)
    G = H.transpose(-1, -2) @ H
    trG = G.diagonal(dim1=-2, dim2=-1).sum(-1)
    if name == "NearlyIncompressibleMooneyRivlin":
        alpha, beta, kappa = p["mu1"], p["mu2"], p["lamb"]
        return alpha * J ** (-2.0 / 3.0) * trC + beta * J ** (-2.0) * trG ** 1.5 + kappa / 2.0 * (J - 1.0) ** 2
    if name == "AnisotropicMooneyRivlin_1":
        mu1, mu2, mu3, lamb = p["mu1"], p["mu2"], p["mu3"], p["lamb"]
        FN = torch.einsum("...ij,...j->...i", F, N)
        HN = torch.einsum("...ij,...j->...i", H, N)
        return mu1 * trC + mu2 * trG - 2.0 * (mu1 + 2.0 * mu2 + mu3) * torch.log(J) + lamb / 2.0 * (J - 1.0) ** 2 \
            + mu3 * (FN * FN).sum(-1) + mu3 * (HN * HN).sum(-1)
    raise NotImplementedError(name)


def potential(xflat, points, elements, Jm, AllGauss, material):
    X = torch.as_tensor(np.asarray(points), dtype=torch.float64)
    el = torch.as_tensor(np.asarray(elements, dtype=np.int64))
    Jmt = torch.as_tensor(np.asarray(Jm), dtype=torch.float64)
    w = torch.as_tensor(np.asarray(AllGauss).ravel(), dtype=torch.float64)
    x = xflat.reshape(-1, 3)
    Xe, xe = X[el], x[el]
    PGX = torch.einsum("iag,eal->egil", Jmt, Xe)        # dX_l/dxi_i
    PGx = torch.einsum("iag,eal->egil", Jmt, xe)
    # F_lj = dx_l/dX_j = sum_i (dx_l/dxi_i)(dxi_i/dX_j)
    F = torch.einsum("egil,egji->eglj", PGx, torch.linalg.inv(PGX))
    dV = w[None, :] * torch.abs(torch.linalg.det(PGX))
    ori = material.get("anisotropic_orientations") if isinstance(material, dict) else \
        getattr(material, "anisotropic_orientations", None)
    N = None if ori is None else torch.as_tensor(np.asarray(ori), dtype=torch.float64)[:, None, :]
    return (dV * strain_energy(F, material, N)).sum()


def gradient_and_hessian(points, elements, Eulerx, Jm, AllGauss, material):
    x0 = torch.as_tensor(np.asarray(Eulerx), dtype=torch.float64).reshape(-1).clone()
    f = lambda xf: potential(xf, points, elements, Jm, AllGauss, material)
    g = torch.func.grad(f)(x0)
    Hs = torch.func.hessian(f)(x0)
    return g.numpy(), Hs.numpy()
