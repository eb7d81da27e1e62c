"""FP64 numpy restatement of the reference's low-level displacement-formulation
assembly -- TEST INFRASTRUCTURE (see oracle/__init__.py).

Every function cites the reference lines it restates.  Element loops are
vectorised over a batch of elements; within an element the arithmetic follows the
reference's order of operations where that order is visible (3x3 adjugate inverse,
dense B, H*B^T then B*(H*B^T), geometric term added after the constitutive term,
ascending-element CSR accumulation).
"""
import numpy as np

# Voigt pairs (Tensor/Numeric.pyx:181-235 _Voigt3; _ConstitutiveStiffnessDF_.h:15-23)
VOIGT = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]


# ----------------------------------------------------------------------- kinematics
def invdet3x3(A):
    """Adjugate inverse and determinant of (..., 3, 3) -- Tensor/_det_inv_.h:149-190."""
    s = A.reshape(A.shape[:-2] + (9,))
    s0, s1, s2, s3, s4, s5, s6, s7, s8 = [s[..., i] for i in range(9)]
    d = np.empty_like(s)
    d[..., 0] = s4 * s8 - s5 * s7
    d[..., 1] = -s1 * s8 + s2 * s7
    d[..., 2] = s1 * s5 - s2 * s4
    d[..., 3] = -s3 * s8 + s5 * s6
    d[..., 4] = s0 * s8 - s2 * s6
    d[..., 5] = -s0 * s5 + s2 * s3
    d[..., 6] = s3 * s7 - s4 * s6
    d[..., 7] = -s0 * s7 + s1 * s6
    d[..., 8] = s0 * s4 - s1 * s3
    det = s0 * d[..., 0] + s1 * d[..., 3] + s2 * d[..., 6]
    inv = d * (1.0 / det)[..., None]
    return inv.reshape(A.shape), det


def kinematic_measures(Jm, AllGauss, X, x, update=1):
    """Per element and quadrature point kinematics -- _KinematicMeasures_.h:64-119.

    Jm: (3, npe, G); AllGauss: (G,) or (G, 1); X, x: (E, npe, 3).
    Returns SpatialGradient (E, G, npe, 3), F (E, G, 3, 3), detJ (E, G):
      ParentGradientX[i][l] = sum_a Jm[i,a,g] X[a,l]         (:88)
      ParentGradientx[i][l] = sum_a Jm[i,a,g] x[a,l]         (:89)
      detJ = w_g |det ParentGradientx| (update) else |det ParentGradientX|   (:94-99)
      MaterialGradient = inv(PGX) @ Jm_g ; spatial = inv(PGx) @ Jm_g          (:101-108)
      F = (MaterialGradient @ x)^T, i.e. F_ij = sum_a x_ai dN_a/dX_j         (:111-113)
    """
    w = np.asarray(AllGauss, dtype=np.float64).ravel()
    PGX = np.einsum("iag,eal->egil", Jm, X)
    PGx = np.einsum("iag,eal->egil", Jm, x)
    iX, detX = invdet3x3(PGX)
    ix, detx = invdet3x3(PGx)
    detJ = w[None, :] * np.abs(detx if update == 1 else detX)
    MG = np.einsum("egjk,kag->egja", iX, Jm)          # (E,G,3,npe)
    SG = np.einsum("egjk,kag->egaj", ix, Jm)          # (E,G,npe,3)
    Ft = np.einsum("egja,eal->egjl", MG, x)           # current_F[j][l]
    F = np.swapaxes(Ft, -1, -2)
    return SG, F, detJ


# ------------------------------------------------------------------------ materials
def _delta():
    return np.eye(3)


def neo_hookean(F, mu, lamb):
    """Cauchy stress and Voigt tangent -- LLDispatch/CythonSource/_NeoHookean_.h:27-51.

    sigma = mu/J (b - I) + lamb (J-1) I
    c_ijkl = (mu/J - lamb(J-1)) (d_ik d_jl + d_il d_jk) + lamb (2J-1) d_ij d_kl
    """
    I = _delta()
    J = np.linalg.det(F)
    b = F @ np.swapaxes(F, -1, -2)
    sig = (mu / J)[..., None, None] * (b - I) + (lamb * (J - 1.0))[..., None, None] * I
    II_ijkl = np.einsum("ij,kl->ijkl", I, I)
    II_ikjl = np.einsum("ik,jl->ijkl", I, I)
    II_iljk = np.einsum("il,jk->ijkl", I, I)
    a1 = (mu / J - lamb * (J - 1.0))[..., None, None, None, None]
    a2 = (lamb * (2.0 * J - 1.0))[..., None, None, None, None]
    c = a1 * (II_ikjl + II_iljk) + a2 * II_ijkl
    return sig, voigt(c)


def mooney_rivlin(F, mu1, mu2, lamb):
    """Cauchy stress and Voigt tangent -- LLDispatch/CythonSource/_MooneyRivlin_.h:30-70.

    sigma = 2mu1/J b + 2mu2/J (tr(b) b - b b) - 2(mu1+2mu2)/J I + lamb(J-1) I
    c_ijkl = 2mu2/J (2 b_ij b_kl - b_ik b_jl - b_il b_jk)
             + (2(mu1+2mu2)/J - lamb(J-1)) (d_ik d_jl + d_il d_jk) + lamb(2J-1) d_ij d_kl
    """
    I = _delta()
    J = np.linalg.det(F)
    b = F @ np.swapaxes(F, -1, -2)
    trb = np.trace(b, axis1=-2, axis2=-1)
    E = lambda v: v[..., None, None]
    sig = E(2.0 * mu1 / J) * b + E(2.0 * mu2 / J) * (E(trb) * b - b @ b) \
        - E(2.0 * (mu1 + 2 * mu2) / J) * I + E(lamb * (J - 1.0)) * I
    II_ijkl = np.einsum("ij,kl->ijkl", I, I)
    II_ikjl = np.einsum("ik,jl->ijkl", I, I)
    II_iljk = np.einsum("il,jk->ijkl", I, I)
    bb_ijkl = np.einsum("...ij,...kl->...ijkl", b, b)
    bb_ikjl = np.einsum("...ik,...jl->...ijkl", b, b)
    bb_iljk = np.einsum("...il,...jk->...ijkl", b, b)
    E4 = lambda v: v[..., None, None, None, None]
    c = E4(2.0 * mu2 / J) * (2.0 * bb_ijkl - bb_ikjl - bb_iljk) \
        + E4(2.0 * (mu1 + 2 * mu2) / J - lamb * (J - 1.0)) * (II_ikjl + II_iljk) \
        + E4(lamb * (2.0 * J - 1.0)) * II_ijkl
    return sig, voigt(c)


def linear_elastic(F, mu, lamb):
    """Cauchy stress and Voigt tangent -- LLDispatch/CythonSource/_LinearElastic_.h:26-52.

    strain = (GradU + GradU^T)/2 with GradU = F - I
    sigma  = 2 mu strain + lamb tr(strain) I
    c_ijkl = mu (d_ik d_jl + d_il d_jk) + lamb d_ij d_kl   (independent of F)
    The low-level assembler drops the geometric stiffness for this law
    (AOT_Assembler.py:84-93).
    """
    I = _delta()
    GradU = F - I
    strain = 0.5 * (GradU + np.swapaxes(GradU, -1, -2))
    tre = np.trace(strain, axis1=-2, axis2=-1)
    sig = 2.0 * mu * strain + (lamb * tre)[..., None, None] * I
    II_ijkl = np.einsum("ij,kl->ijkl", I, I)
    II_ikjl = np.einsum("ik,jl->ijkl", I, I)
    II_iljk = np.einsum("il,jk->ijkl", I, I)
    c = mu * (II_ikjl + II_iljk) + lamb * II_ijkl
    H = np.broadcast_to(voigt(c), F.shape[:-2] + (6, 6)).copy()
    return sig, H


def nearly_incompressible_mooney_rivlin(F, alpha, beta, kappa):
    """Cauchy stress and Voigt tangent --
    LLDispatch/CythonSource/_NearlyIncompressibleMooneyRivlin_.h:28-87.

    H = cof(F), b = F F^T, g = H H^T; c0 = J^(-5/3), c1 = sqrt(tr g), c2 = 1/c1,
    c3 = tr(g) c1, c4 = J^-3:
    sigma = 2 alpha c0 b - 2/3 alpha c0 tr(b) I + beta c4 c3 I - 3 beta c4 c1 g
            + kappa (J-1) I
    c_ijkl = -4/3 alpha c0 (b_ij d_kl + d_ij b_kl) + 4/9 alpha c0 tr(b) d_ij d_kl
             + 2/3 alpha c0 tr(b) (d_ik d_jl + d_il d_jk)
             + beta c4 c3 (d_ij d_kl - d_ik d_jl - d_il d_jk)
             - 3 beta c4 c1 (g_ij d_kl + d_ij g_kl)
             + 3 beta c4 c1 (g_ik d_jl + g_il d_jk + d_ik g_jl + d_il g_jk)
             + 3 beta c4 c2 g_ij g_kl
             + kappa (2J-1) d_ij d_kl - kappa (J-1) (d_ik d_jl + d_il d_jk)
    The low-level assembler constructs it from (mu1, mu2, lamb) as (alpha, beta, kappa)
    (AOT_Assembler.py:69-70, 114-115), which is how the dispatcher below calls it.
    """
    I = _delta()
    J = np.linalg.det(F)
    Hc = J[..., None, None] * np.swapaxes(np.linalg.inv(F), -1, -2)   # cofactor(F)
    b = F @ np.swapaxes(F, -1, -2)
    g = Hc @ np.swapaxes(Hc, -1, -2)
    trb = np.trace(b, axis1=-2, axis2=-1)
    trg = np.trace(g, axis1=-2, axis2=-1)
    c0 = J ** (-5.0 / 3.0)
    c1 = np.sqrt(trg)
    c2 = 1.0 / c1
    c3 = trg * c1
    c4 = 1.0 / (J * J * J)
    E = lambda v: v[..., None, None]
    sig = E(2.0 * alpha * c0) * b - E(2.0 / 3.0 * alpha * c0 * trb) * I + E(beta * c4 * c3) * I \
        - E(3.0 * beta * c4 * c1) * g + E(kappa * (J - 1.0)) * I
    E4 = lambda v: v[..., None, None, None, None]
    Ib = np.broadcast_to(I, b.shape)
    II_ijkl = np.einsum("ij,kl->ijkl", I, I)
    II_ikjl = np.einsum("ik,jl->ijkl", I, I)
    II_iljk = np.einsum("il,jk->ijkl", I, I)
    bI = np.einsum("...ij,...kl->...ijkl", b, Ib) + np.einsum("...ij,...kl->...ijkl", Ib, b)
    gI = np.einsum("...ij,...kl->...ijkl", g, Ib) + np.einsum("...ij,...kl->...ijkl", Ib, g)
    # gI_ikjl + gI_iljk + Ig_ikjl + Ig_iljk
    gIp = np.einsum("...ik,...jl->...ijkl", g, Ib) + np.einsum("...il,...jk->...ijkl", g, Ib) \
        + np.einsum("...ik,...jl->...ijkl", Ib, g) + np.einsum("...il,...jk->...ijkl", Ib, g)
    gg = np.einsum("...ij,...kl->...ijkl", g, g)
    c = E4(-4.0 / 3.0 * alpha * c0) * bI + E4(4.0 * alpha / 9.0 * c0 * trb) * II_ijkl \
        + E4(2.0 / 3.0 * alpha * c0 * trb) * (II_ikjl + II_iljk) \
        + E4(beta * c4 * c3) * (II_ijkl - II_ikjl - II_iljk) \
        - E4(3.0 * beta * c4 * c1) * gI + E4(3.0 * beta * c4 * c1) * gIp + E4(3.0 * beta * c4 * c2) * gg \
        + E4(kappa * (2.0 * J - 1.0)) * II_ijkl - E4(kappa * (J - 1.0)) * (II_ikjl + II_iljk)
    return sig, voigt(c)


def anisotropic_mooney_rivlin_1(F, N, mu1, mu2, mu3, lamb):
    """Cauchy stress and Voigt tangent --
    LLDispatch/CythonSource/_AnisotropicMooneyRivlin_1_.h:30-95; N is the element's
    fibre direction (..., 3) (the assembler passes anisotropic_orientations[elem],
    AOT_Assembler.py:78-82).

    H = cof(F), FN = F N, HN = H N:
    sigma = 2mu1/J b + 2mu2/J (tr(b) b - b b) - 2(mu1+2mu2+mu3)/J I + lamb(J-1) I
            + 2mu3/J FN FN^T + 2mu3/J (HN.HN) I - 2mu3/J HN HN^T
    c_ijkl = 2mu2/J (2 b_ij b_kl - b_ik b_jl - b_il b_jk)
             + (2(mu1+2mu2+mu3)/J - lamb(J-1)) (d_ik d_jl + d_il d_jk) + lamb(2J-1) d_ij d_kl
             - 4mu3/J (d_ij n_k n_l + n_i n_j d_kl)
             + 2mu3/J (HN.HN) (2 d_ij d_kl - d_ik d_jl - d_il d_jk)
             + 2mu3/J (d_ik n_j n_l + d_il n_j n_k + d_jl n_i n_k + d_jk n_i n_l),  n = HN
    """
    I = _delta()
    J = np.linalg.det(F)
    Hc = J[..., None, None] * np.swapaxes(np.linalg.inv(F), -1, -2)
    b = F @ np.swapaxes(F, -1, -2)
    N = np.broadcast_to(N, F.shape[:-1])
    FN = np.einsum("...ij,...j->...i", F, N)
    HN = np.einsum("...ij,...j->...i", Hc, N)
    oFN = np.einsum("...i,...j->...ij", FN, FN)
    oHN = np.einsum("...i,...j->...ij", HN, HN)
    iHN = np.einsum("...i,...i->...", HN, HN)
    trb = np.trace(b, axis1=-2, axis2=-1)
    E = lambda v: v[..., None, None]
    sig = E(2.0 * mu1 / J) * b + E(2.0 * mu2 / J) * (E(trb) * b - b @ b) \
        - E(2.0 * (mu1 + 2.0 * mu2 + mu3) / J) * I + E(lamb * (J - 1.0)) * I \
        + E(2.0 * mu3 / J) * oFN + E(2.0 * mu3 / J * iHN) * I - E(2.0 * mu3 / J) * oHN
    E4 = lambda v: v[..., None, None, None, None]
    Ib = np.broadcast_to(I, b.shape)
    II_ijkl = np.einsum("ij,kl->ijkl", I, I)
    II_ikjl = np.einsum("ik,jl->ijkl", I, I)
    II_iljk = np.einsum("il,jk->ijkl", I, I)
    bb = np.einsum("...ij,...kl->...ijkl", b, b)
    bb_ikjl = np.einsum("...ik,...jl->...ijkl", b, b)
    bb_iljk = np.einsum("...il,...jk->...ijkl", b, b)
    IHN = np.einsum("...ij,...kl->...ijkl", Ib, oHN)
    HNI = np.einsum("...ij,...kl->...ijkl", oHN, Ib)
    # IHN_ikjl + IHN_iljk + IHN_jlik + IHN_jkil
    IHNp = np.einsum("...ik,...jl->...ijkl", Ib, oHN) + np.einsum("...il,...jk->...ijkl", Ib, oHN) \
        + np.einsum("...jl,...ik->...ijkl", Ib, oHN) + np.einsum("...jk,...il->...ijkl", Ib, oHN)
    c = E4(2.0 * mu2 / J) * (2.0 * bb - bb_ikjl - bb_iljk) \
        + E4(2.0 * (mu1 + 2.0 * mu2 + mu3) / J - lamb * (J - 1.0)) * (II_ikjl + II_iljk) \
        + E4(lamb * (2.0 * J - 1.0)) * II_ijkl - E4(4.0 * mu3 / J) * (IHN + HNI) \
        + E4(2.0 * mu3 / J * iHN) * (2.0 * II_ijkl - II_ikjl - II_iljk) + E4(2.0 * mu3 / J) * IHNp
    return sig, voigt(c)


def voigt(c):
    """4th-order (..., 3,3,3,3) -> (..., 6, 6): H[A,B] = (c_ijkl + c_ijlk)/2 with
    A=(i,j), B=(k,l) in Voigt order (00,11,22,01,02,12) -- Tensor/Numeric.pyx:181-218."""
    H = np.empty(c.shape[:-4] + (6, 6))
    for A, (i, j) in enumerate(VOIGT):
        for B, (k, l) in enumerate(VOIGT):
            H[..., A, B] = 0.5 * (c[..., i, j, k, l] + c[..., i, j, l, k])
    return H


def material_name(material):
    return type(material).__name__ if not isinstance(material, dict) else material["name"]


def has_geometric_stiffness(material):
    """The low-level LinearElastic assembler is generated without the geometric term
    (AOT_Assembler.py:84-93); every other law keeps it."""
    return material_name(material) != "LinearElastic"


def kinetic_measures(material, F, orientation=None):
    """Dispatch on the material name (as _LowLevelAssembly_.py:47-60), with the
    constants the generated assemblers construct each law from
    (AOT_Assembler.py:64-72).  orientation: (E, 3) fibre directions of the elements of
    F (E, G, 3, 3), for AnisotropicMooneyRivlin_1."""
    name = material_name(material)
    p = material if isinstance(material, dict) else material.__dict__
    if name == "NeoHookean":
        return neo_hookean(F, p["mu"], p["lamb"])
    if name == "MooneyRivlin":
        return mooney_rivlin(F, p["mu1"], p["mu2"], p["lamb"])
    if name == "LinearElastic":
        return linear_elastic(F, p["mu"], p["lamb"])
    if name == "NearlyIncompressibleMooneyRivlin":
        return nearly_incompressible_mooney_rivlin(F, p["mu1"], p["mu2"], p["lamb"])
    if name == "AnisotropicMooneyRivlin_1":
        N = np.asarray(orientation, dtype=np.float64)
        if F.ndim == 4:
            N = N[:, None, :]
        return anisotropic_mooney_rivlin_1(F, N, p["mu1"], p["mu2"], p["mu3"], p["lamb"])
    raise NotImplementedError(name)


# ------------------------------------------------------------- element integrands
def constitutive_B(SG):
    """B (..., npe*3, 6) from spatial gradients (..., npe, 3) -- FillConstitutiveB_
    (_ConstitutiveStiffnessDF_.h:32-59): rows 3a+d, columns Voigt."""
    sh = SG.shape[:-2]
    npe = SG.shape[-2]
    B = np.zeros(sh + (npe, 3, 6))
    a0, a1, a2 = SG[..., 0], SG[..., 1], SG[..., 2]
    B[..., 0, 0] = a0
    B[..., 1, 1] = a1
    B[..., 2, 2] = a2
    B[..., 1, 5] = a2
    B[..., 2, 5] = a1
    B[..., 0, 4] = a2
    B[..., 2, 4] = a0
    B[..., 0, 3] = a1
    B[..., 1, 3] = a0
    return B.reshape(sh + (npe * 3, 6))


def total_traction(sig):
    """[s00, s11, s22, s01, s02, s12] -- GetTotalTraction_ (_ConstitutiveStiffnessDF_.h:15-23)."""
    s = sig.reshape(sig.shape[:-2] + (9,))
    return np.stack([s[..., 0], s[..., 4], s[..., 8], s[..., 1], s[..., 2], s[..., 5]], axis=-1)


def local_stiffness_traction(SG, sig, H, detJ, requires_geometry_update=1, geometric=True):
    """Element stiffness (E, n, n) and traction (E, n).

    Constitutive: sum_g detJ_g * B (H B^T)  (_ConstitutiveStiffnessDF_.h:109-134)
    Traction:     sum_g detJ_g * B s        (:136-148)
    Geometric:    K[3a+d, 3b+d] += detJ_g * gA . sigma . gB  (_GeometricStiffness_.h:61-100)
    Sum:          stiffness += geometric (_LowLevelAssemblyDF_.h:128-130)
    geometric=False: the term is left out (the LinearElastic assembler,
    AOT_Assembler.py:84-93).
    """
    B = constitutive_B(SG)                                  # (E,G,n,6)
    E, G, n = B.shape[0], B.shape[1], B.shape[2]
    K = np.zeros((E, n, n))
    for g in range(G):                                      # per quadrature point, as the reference
        HBT = np.einsum("eAB,emB->eAm", H[:, g], B[:, g])   # (E,6,n)
        BDB = np.matmul(B[:, g], HBT)                       # (E,n,n)
        K += BDB * detJ[:, g, None, None]
    t = np.zeros(K.shape[:2])
    if requires_geometry_update == 1:
        s = total_traction(sig)
        t = np.einsum("egnA,egA,eg->en", B, s, detJ)
    if not geometric:
        return K, t
    # geometric stiffness with symmetric sigma (uses s00,s01,s02,s11,s12,s22 as the C code)
    sg = sig.copy()
    sg[..., 1, 0] = sg[..., 0, 1]
    sg[..., 2, 0] = sg[..., 0, 2]
    sg[..., 2, 1] = sg[..., 1, 2]
    dum = np.einsum("egai,egij,egbj,eg->eab", SG, sg, SG, detJ)
    E, npe = dum.shape[:2]
    Kg = np.zeros((E, npe, 3, npe, 3))
    for d in range(3):
        Kg[:, :, d, :, d] = dum
    K = K + Kg.reshape(E, 3 * npe, 3 * npe)
    return K, t


def element_matrices(points, elements, Eulerx, Jm, AllGauss, material, requires_geometry_update=1,
                     elem_slice=None):
    """(K_e, t_e) for a slice of elements: gather -> kinematics -> kinetics ->
    integrands, i.e. the body of the element loop of _LowLevelAssemblyDF_.h:69-130."""
    el = np.asarray(elements, dtype=np.int64)
    sl = elem_slice if elem_slice is not None else slice(None)
    el = el[sl]
    X = np.asarray(points)[el]
    x = np.asarray(Eulerx)[el]
    SG, F, detJ = kinematic_measures(Jm, AllGauss, X, x, requires_geometry_update)
    sig, H = kinetic_measures(material, F, element_orientations(material, sl))
    return local_stiffness_traction(SG, sig, H, detJ, requires_geometry_update, has_geometric_stiffness(material))


def element_orientations(material, sl):
    """Rows of anisotropic_orientations (nelem, 3) for an element slice, or None."""
    ori = material.get("anisotropic_orientations") if isinstance(material, dict) else \
        getattr(material, "anisotropic_orientations", None)
    return None if ori is None else np.asarray(ori, dtype=np.float64)[sl]


# ------------------------------------------------------------------ sparse pattern
def sparsity_pattern(elements, nnode, nvar=3):
    """CSR indptr/indices (int32) -- ComputeSparsityPattern.pyx:44-90 and
    _ComputeSparsityPattern_ (ComputeSparsityPattern.h:22-62): for every node the
    sorted unique set of nodes sharing an element with it; its nvar rows each list
    nvar*nbr+l for every neighbour nbr and l in [0, nvar)."""
    el = np.asarray(elements, dtype=np.int64)
    nelem, npe = el.shape
    flat = el.ravel()
    order = np.argsort(flat, kind="stable")
    elem_of = order // npe
    starts = np.searchsorted(flat[order], np.arange(nnode + 1))
    counts = np.zeros(nnode, dtype=np.int64)
    nbrs = []
    for a in range(nnode):
        es = elem_of[starts[a]:starts[a + 1]]
        nb = np.unique(el[es].ravel())
        nbrs.append(nb)
        counts[a] = nb.shape[0]
    indices = np.concatenate([
        np.tile((nvar * nb[:, None] + np.arange(nvar)[None, :]).ravel(), nvar) for nb in nbrs
    ]).astype(np.int32)
    rowlen = np.repeat(counts * nvar, nvar)
    indptr = np.zeros(nnode * nvar + 1, dtype=np.int64)
    np.cumsum(rowlen, out=indptr[1:])
    return indptr.astype(np.int32), indices


def data_global_indices(indptr, indices, elements, nvar=3):
    """Position in CSR data of every local entry (e, i, j) of the element matrices --
    _ComputeDataIndices_ (ComputeSparsityPattern.h:67-122), by binary search of the
    global column in the row's sorted column list."""
    el = np.asarray(elements, dtype=np.int64)
    nelem, npe = el.shape
    ndof = nvar * npe
    dofs = (nvar * el[:, :, None] + np.arange(nvar)[None, None, :]).reshape(nelem, ndof)
    out = np.empty((nelem, ndof, ndof), dtype=np.int64)
    indptr = np.asarray(indptr, dtype=np.int64)
    for e in range(nelem):
        for i in range(ndof):
            r = dofs[e, i]
            row = indices[indptr[r]:indptr[r + 1]]
            out[e, i] = indptr[r] + np.searchsorted(row, dofs[e])
    return out


def assemble_csr(points, elements, Eulerx, Jm, AllGauss, material, indptr, indices,
                 requires_geometry_update=1, chunk=2048, return_local=False):
    """Global CSR data and residual T -- the CSR branch of fill_global_data
    (_MassIntegrand_.h:115-162 -> SparseAssemblyNativeCSR_, SparseAssemblyNative.h:32-45)
    and the traction scatter (_LowLevelAssemblyDF_.h:153-161), accumulating in
    ascending element order."""
    el = np.asarray(elements, dtype=np.int64)
    nelem, npe = el.shape
    nnode = np.asarray(points).shape[0]
    V = np.zeros(np.asarray(indices).shape[0])
    T = np.zeros(3 * nnode)
    locals_ = []
    for s in range(0, nelem, chunk):
        sl = slice(s, min(nelem, s + chunk))
        K, t = element_matrices(points, el, Eulerx, Jm, AllGauss, material, requires_geometry_update, sl)
        gidx = data_global_indices(indptr, indices, el[sl])
        for k in range(K.shape[0]):
            # ascending element order; within an element every position is distinct
            V[gidx[k].ravel()] += K[k].ravel()
            dof = (3 * el[s + k][:, None] + np.arange(3)[None, :]).ravel()
            T[dof] += t[k]
        if return_local:
            locals_.append((K, t))
    if return_local:
        return V, T, locals_
    return V, T


def residual_only(points, elements, Eulerx, Jm, AllGauss, material, chunk=4096):
    """Matrix-free internal force -- the mechanics branch of
    _GlobalAssemblyExplicit_DF_DPF_ (_LowLevelAssemblyExplicit_DF_DPF_.h:459-726):
    T_a += sum_g detJ_g sigma_g grad_x N_a, scattered in ascending element order."""
    el = np.asarray(elements, dtype=np.int64)
    nelem, npe = el.shape
    T = np.zeros(3 * np.asarray(points).shape[0])
    for s in range(0, nelem, chunk):
        sl = slice(s, min(nelem, s + chunk))
        X = np.asarray(points)[el[sl]]
        x = np.asarray(Eulerx)[el[sl]]
        SG, F, detJ = kinematic_measures(Jm, AllGauss, X, x, 1)
        sig, _ = kinetic_measures(material, F, element_orientations(material, sl))
        t = np.einsum("egaj,egij,eg->eai", SG, sig, detJ).reshape(-1, 3 * npe)
        for k in range(t.shape[0]):
            dof = (3 * el[s + k][:, None] + np.arange(3)[None, :]).ravel()
            T[dof] += t[k]
    return T
