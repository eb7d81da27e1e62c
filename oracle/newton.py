"""Incremental Newton-Raphson restated on the CPU -- TEST INFRASTRUCTURE
(see oracle/__init__.py).

Follows Florence/Solver/FEMSolver.py StaticSolver (:841-929) and NewtonRaphson
(:932-1044) with the reference's reduced system (BoundaryCondition.py:842-933:
K[in, in], predictor K[in, out] u_D * LoadFactor) and a sparse direct solve
(scipy.sparse.linalg.spsolve, SuperLU, as LinearSolver.py:253-270 falls back to when
UMFPACK is absent).  Assembly is the C restatement (oracle/df_ref.c) when built, else
the numpy restatement.
"""
import ctypes
import os

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import df_numpy as O

_HERE = os.path.dirname(os.path.abspath(__file__))


def _c_lib():
    so = os.path.join(_HERE, "liboracle_df.so")
    if not os.path.exists(so):
        return None
    lib = ctypes.CDLL(so)
    lib.oracle_df_assemble_ex.restype = ctypes.c_int
    return lib


_IDS = {"NeoHookean": 1, "MooneyRivlin": 2, "LinearElastic": 3, "NearlyIncompressibleMooneyRivlin": 4,
        "AnisotropicMooneyRivlin_1": 5}


def make_assembler(mesh, fs, material, indptr, indices):
    """x -> (V, T): the C restatement (liboracle_df.so) when built, else the numpy one."""
    lib = _c_lib()
    el = np.ascontiguousarray(mesh.elements, dtype=np.int64)
    pts = np.ascontiguousarray(mesh.points, dtype=np.float64)
    Jm, w = np.ascontiguousarray(fs.Jm), np.ascontiguousarray(fs.AllGauss.ravel())
    ip, ix = np.ascontiguousarray(indptr, dtype=np.int32), np.ascontiguousarray(indices, dtype=np.int32)
    name = type(material).__name__
    f = lambda a: 0.0 if getattr(material, a, None) is None else float(getattr(material, a))
    prm = np.array([f("mu"), f("mu1"), f("mu2"), f("lamb"), f("mu3")], dtype=np.float64)
    ori = getattr(material, "anisotropic_orientations", None)
    ori = None if ori is None else np.ascontiguousarray(ori, dtype=np.float64)
    mid = _IDS[name]
    P = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)

    def assemble(x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        if lib is None:
            return O.assemble_csr(pts, el, x, Jm, w, material, ip, ix)
        V = np.zeros(ix.size)
        T = np.zeros(3 * pts.shape[0])
        rc = lib.oracle_df_assemble_ex(P(pts), P(el), P(x), P(Jm), P(w), ctypes.c_int64(w.size),
                                       ctypes.c_int64(el.shape[1]), ctypes.c_int(mid), P(prm), P(ori),
                                       ctypes.c_int(1), P(ip), P(ix), P(V), P(T), ctypes.c_int64(0),
                                       ctypes.c_int64(el.shape[0]))
        if rc != 0:
            raise RuntimeError("oracle assembly failed")
        return V, T

    return assemble


def solve_static(mesh, fs, material, dirichlet_flags, nincr, tol=1e-8, maxit=50, break_at_stagnation=True,
                 record=None):
    """Returns (TotalDisp (nnode, 3, nincr), NRConvergence dict, iterations per increment).

    record (optional dict) receives per increment: "first_iterate" (the displacement after
    the first Newton update, i.e. the predictor step), "abs_residual" (the absolute
    ||R_free|| after every iteration) and "predictor_residual" (||R_free|| of the
    predictor, the reference's NormForces at increment 0)."""
    nnode = mesh.points.shape[0]
    ndof = 3 * nnode
    indptr, indices = O.sparsity_pattern(mesh.elements, nnode)
    assemble = make_assembler(mesh, fs, material, indptr, indices)
    flat = np.asarray(dirichlet_flags, dtype=np.float64).ravel()
    cols_out = np.arange(flat.size)[~np.isnan(flat)]
    applied = flat[~np.isnan(flat)]
    cols_in = np.delete(np.arange(ndof), cols_out)
    sol_tol = 10.0 * tol
    Eulerx = np.copy(mesh.points).ravel()
    V, T = assemble(Eulerx.reshape(-1, 3))
    K = sp.csr_matrix((V, indices, indptr), shape=(ndof, ndof))
    NodalForces = np.zeros(ndof)
    TotalDisp = np.zeros((nnode, 3, nincr))
    NR = {"Increment_" + str(i): [] for i in range(nincr)}
    iters = []
    Residual = np.zeros(ndof)
    NormForces = None
    for inc in range(nincr):
        LF = 1.0 / nincr
        nz = ~np.isclose(applied, 0.0)
        Residual = np.zeros(ndof)
        Residual[cols_in] = K[cols_in, :][:, cols_out[nz]] @ (applied[nz] * LF)
        if inc == 0:
            NormForces = np.linalg.norm(Residual)
            if np.isclose(NormForces, 0.0):
                NormForces = 1e-14
        norm_residual = np.linalg.norm(Residual) / NormForces
        if record is not None:
            record.setdefault("predictor_residual", []).append(np.linalg.norm(Residual))
            record.setdefault("abs_residual", []).append([])
        Eulerx[cols_out] += LF * applied
        Iter = 0
        hist = []
        while norm_residual > tol or Iter == 0:
            K_b = K[cols_in, :][:, cols_in].tocsc()
            sol = spla.spsolve(K_b, -Residual[cols_in])
            dU = np.zeros(ndof)
            dU[cols_in] = sol
            Eulerx += dU
            V, T = assemble(Eulerx.reshape(-1, 3))
            K = sp.csr_matrix((V, indices, indptr), shape=(ndof, ndof))
            Residual[cols_in] = T[cols_in] - NodalForces[cols_in]
            absn = np.linalg.norm(Residual[cols_in])
            if record is not None:
                record["abs_residual"][-1].append(absn)
                if Iter == 0:
                    record.setdefault("first_iterate", []).append(Eulerx.reshape(-1, 3) - mesh.points)
            if Iter == 0:
                NormForces = absn if not np.isclose(absn, 0.0) else 1e-14
            norm_residual = abs(absn / NormForces)
            NR["Increment_" + str(inc)].append(norm_residual)
            if abs(absn) < tol:
                break
            if Iter > 0 and np.linalg.norm(dU) <= sol_tol:
                break
            Iter += 1
            if Iter == maxit or np.isnan(norm_residual) or norm_residual > 1e6:
                raise RuntimeError("oracle Newton failed at increment {}".format(inc))
            if break_at_stagnation:
                hist.append(norm_residual)
                if Iter >= 6 and np.mean(hist) < 1. and absn < 0.001:
                    break
        iters.append(Iter + 1)
        TotalDisp[:, :, inc] = Eulerx.reshape(-1, 3) - mesh.points
    return TotalDisp, NR, iters
