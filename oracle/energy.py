"""Independent energy oracle -- TEST INFRASTRUCTURE (see oracle/__init__.py).

Total potential energy of a small mesh in torch float64, with derivatives by
autograd:

    Pi(x) = sum_e sum_g w_g |det(dX/dxi)| W(F_g),   F = dx/dX

with the strain energies stated by the reference's material docstrings and
``InternalEnergy`` methods:
  Neo-Hookean  W = mu/2 (tr C - 3) - mu ln J + lamb/2 (J-1)^2        (NeoHookean.py:70-84)
  Mooney-Rivlin W = mu1 (tr C - 3) + mu2 (tr G - 3) - 2(mu1+2mu2) ln J + lamb/2 (J-1)^2,
                G = H^T H, H = J F^{-T}                              (MooneyRivlin.py:80-95)

dPi/dx must equal the assembled internal force T and d2Pi/dx2 the assembled
tangent K (updated-Lagrangian integration is exact per quadrature point for
det(dx/dxi) > 0), which pins the kinematics, the material laws, the Voigt
convention, B, the geometric term and the scatter at once.
"""
import numpy as np
import torch


def strain_energy(F, material):
    name = type(material).__name__ if not isinstance(material, dict) else material["name"]
    p = material if isinstance(material, dict) else material.__dict__
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
    return (dV * strain_energy(F, material)).sum()


def gradient_and_hessian(points, elements, Eulerx, Jm, AllGauss, material):
    x0 = torch.as_tensor(np.asarray(Eulerx), dtype=torch.float64).reshape(-1).clone()
    f = lambda xf: potential(xf, points, elements, Jm, AllGauss, material)
    g = torch.func.grad(f)(x0)
    Hs = torch.func.hessian(f)(x0)
    return g.numpy(), Hs.numpy()
