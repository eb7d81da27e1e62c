"""Mesh / material builders shared by the tests (workloads of the assembly path)."""
import numpy as np

from florence_amd.function_space import FunctionSpace
from florence_amd.quadrature import QuadratureRule
from florence_amd.workloads import cube_hex_p4, cube_mesh, cube_tet_p2  # noqa: F401


def function_space(mesh, p, qtype="gauss", norder=None):
    if norder is None:
        norder = 2 * (p - 1) if mesh.element_type == "tet" else p + 1
        norder = max(norder, 1)
    q = QuadratureRule(qtype=qtype, norder=norder, mesh_type=mesh.element_type)
    return FunctionSpace(mesh, q, p=p)


def tet_p2_cube(n=8, seed=0):
    """Unit cube, n^3 cubes x 6 tets, p=2, degree-4 (11 point) rule, jitter +-0.05h, NH(1, 10)."""
    return cube_tet_p2(n, seed)


def hex_p4_cube(n=24, seed=0):
    """Unit cube, n^3 hexes, p=4, 6x6x6 Gauss-Lobatto rule, jitter +-0.03h, MR(0.5, 0.5, 20)."""
    return cube_hex_p4(n, seed)


def deformed(points, amp=0.05, seed=1):
    """A smooth finite deformation of the reference coordinates."""
    X = np.asarray(points)
    rng = np.random.default_rng(seed)
    A = np.eye(3) + amp * rng.standard_normal((3, 3))
    return X @ A.T + amp * np.sin(np.pi * X[:, [1, 2, 0]])


def oracle_assemble_many(items, threads=8):
    """FP64 oracle (C restatement of _GlobalAssemblyDF_, oracle/df_ref.c) for several
    meshes at once: items = [(mesh, function_space, material, x, indptr, indices)];
    returns [(V, T)].  The C calls release the GIL, so meshes run on parallel threads."""
    from concurrent.futures import ThreadPoolExecutor

    from oracle.newton import make_assembler

    def one(it):
        mesh, fs, mat, x, ip, ix = it
        return make_assembler(mesh, fs, mat, ip, ix)(x)

    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(one, items))


class _Shifted(object):
    def __init__(self, mesh, c):
        self.points = np.asarray(mesh.points) + c
        self.elements = mesh.elements


# translation of the second oracle evaluation (a quarter of the unit cube's size)
ORACLE_SHIFT = np.array([0.25, -0.5, 0.75])


def oracle_with_spread(items, threads=8):
    """The oracle of ``items`` (as oracle_assemble_many) and its own FP64 spread: the
    same assembly on the mesh and its deformation translated by ORACLE_SHIFT.  K and T
    are translation invariant, so |V - V_shift|, |T - T_shift| measure the rounding of the
    reference algorithm itself (its parent gradients sum absolute coordinates).  Returns
    [(V, T, max|V - V_shift|, max|T - T_shift|)]."""
    shifted = [(_Shifted(m, ORACLE_SHIFT), fs, mat, x + ORACLE_SHIFT, ip, ix) for m, fs, mat, x, ip, ix in items]
    res = oracle_assemble_many(list(items) + shifted, threads)
    n = len(items)
    return [(V, T, float(np.abs(V - Vs).max()), float(np.abs(T - Ts).max()))
            for (V, T), (Vs, Ts) in zip(res[:n], res[n:])]


def within_oracle(V, T, Vr, Tr, dV, dT, rtol=1e-12):
    """K and T against the oracle: within rtol relative (inf-norm), or within twice the
    oracle's own translation spread where that is larger.  Returns (ok, eK, eT, bound_K,
    bound_T) with the errors and bounds relative to max|Vr|, max|Tr|."""
    sK, sT = np.abs(Vr).max(), np.abs(Tr).max()
    eK, eT = np.abs(V - Vr).max() / sK, np.abs(T - Tr).max() / sT
    bK, bT = max(rtol, 2.0 * dV / sK), max(rtol, 2.0 * dT / sT)
    return eK <= bK and eT <= bT, eK, eT, bK, bT
