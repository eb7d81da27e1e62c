"""p-multigrid linear solver of the Newton steps (no reference counterpart: the
reference solves every step with a sparse direct solver, Solver/LinearSolver.py:208-300).

``MultigridSolver`` has the ``Solve(dm, V, free_mask, b, x)`` interface of
``solver.LinearSolver`` and runs ``fl_mg_pcg``: flexible CG on the masked assembled
tangent, preconditioned by one V-cycle over nested nodal spaces on the mesh's own
elements:

* hex p=4 -> p=2 lattice -> vertices; hex p=2, p=3 -> vertices; tet p>=2 -> vertices;
* coarse operators are Galerkin products P^T K P, formed on the device from the point
  records of the assembly that produced V (``fl_mg_galerkin``), so each Newton step's
  coarse operators follow its tangent.  They are formed element by element (each
  element's own interpolation of the coarse space), which is P^T K P wherever the
  coarse space is nested in the fine one; on the reference's 6-tet split, whose
  neighbouring tets cut a shared cube face along different diagonals, the vertex space
  is not nested, and the coarse operator is that element-by-element sum (still SPD; the
  flexible CG tolerates the mismatch with P);
* transfers are the nodal interpolation of the coarse spaces (a coarse basis function
  is a fine-space function), built here once per mesh as per-node CSR lists.

Host set-up (per mesh): the coarse nodes are the fine nodes at the coarse lattice
positions of each element (found from the local nodes' reference coordinates), their
basis gradients at the fine quadrature points and the transfer weights come from a
Vandermonde solve in the monomial basis of the coarse space, and the coarse
patterns are built by the library's fl_pattern_* calls.
"""
import ctypes

import numpy as np

from . import _lib
from .bases import FeketePointsTet, GaussLobattoPointsHex, HexBases, TetBases

__all__ = ["MultigridSolver", "PMultigridHierarchy", "hierarchy_tables", "local_node_coordinates", "coarse_degrees",
           "matrix_free_operator", "mf_apply"]


class FlMgLevel(ctypes.Structure):
    _fields_ = [("nnode", ctypes.c_int64), ("indptr", ctypes.c_void_p), ("node_nbr_ptr", ctypes.c_void_p),
                ("node_nbrs", ctypes.c_void_p), ("V", ctypes.c_void_p), ("free_mask", ctypes.c_void_p),
                ("P_ptr", ctypes.c_void_p), ("P_idx", ctypes.c_void_p), ("P_w", ctypes.c_void_p),
                ("R_ptr", ctypes.c_void_p), ("R_idx", ctypes.c_void_p), ("R_w", ctypes.c_void_p)]


class FlMgMatrixFree(ctypes.Structure):
    _fields_ = [("df_workspace", ctypes.c_void_p), ("df_workspace_bytes", ctypes.c_size_t),
                ("records", ctypes.c_int32), ("tensor_basis", ctypes.c_void_p), ("elements", ctypes.c_void_p),
                ("nelem", ctypes.c_int64), ("npe", ctypes.c_int32), ("ngauss", ctypes.c_int32),
                ("nnode", ctypes.c_int64), ("node_elem_ptr", ctypes.c_void_p), ("node_elem", ctypes.c_void_p),
                ("element_vectors", ctypes.c_void_p), ("records32", ctypes.c_void_p),
                ("records16", ctypes.c_void_p)]


def matrix_free_operator(dm, precision="fp64", element_vectors=None, records32=None, records16=None):
    """The last assembly's tangent on ``dm`` in the matrix-free form of fl_mf_apply
    (tensor-product hexes whose assembly left C point records), or None where that form
    does not exist.  Returns (descriptor, element-vector buffer).  ``records32``: a float32
    buffer of nelem * 54 * ngauss for fl_mg_pcg's FP32 copy of the C tensors."""
    import torch
    lib = _lib.load()
    if dm.tensor_basis is None:
        return None, None
    rec = int(lib.fl_df_records_form(_lib.PRECISIONS[precision], _lib.ptr(dm.tensor_basis), dm.npe, dm.ngauss))
    if rec != _lib.FL_RECORDS_C:
        return None, None
    n = dm.nelem * 3 * dm.npe
    if element_vectors is None or element_vectors.numel() < n or element_vectors.device != dm.device:
        element_vectors = torch.empty(n, dtype=torch.float64, device=dm.device)
    ws = dm.workspace()
    op = FlMgMatrixFree(ws.data_ptr(), ws.numel(), rec, dm.tensor_basis.data_ptr(), dm.elements.data_ptr(),
                        dm.nelem, dm.npe, dm.ngauss, dm.nnode, dm.node_elem_ptr.data_ptr(), dm.node_elem.data_ptr(),
                        element_vectors.data_ptr(), None if records32 is None else records32.data_ptr(),
                        None if records16 is None else records16.data_ptr())
    return op, element_vectors


def mf_apply(dm, free_mask, x, y, precision="fp64"):
    """y = M K M x of the last assembly on ``dm`` without its CSR (fl_mf_apply)."""
    op, buf = matrix_free_operator(dm, precision)
    if op is None:
        raise _lib.FlorenceAMDError("no matrix-free form for this mesh and precision")
    _lib.check(_lib.load().fl_mf_apply(ctypes.byref(op), _lib.ptr(free_mask), _lib.ptr(x), _lib.ptr(y),
                                       _lib.stream_ptr()), "fl_mf_apply")
    return y


def local_node_coordinates(element_type, npe):
    """Reference coordinates (npe, 3) of the local nodes, checked against the bases
    (N_a(xi_b) = delta_ab)."""
    if element_type == "hex":
        p1 = int(round(npe ** (1.0 / 3.0)))
        C = p1 - 2
        xi = GaussLobattoPointsHex(C)
        N, _ = HexBases(C, xi)
    elif element_type == "tet":
        p = {4: 1, 10: 2, 20: 3, 35: 4}[npe]
        C = p - 1
        xi = FeketePointsTet(C)
        N, _ = TetBases(C, xi)
    else:
        raise NotImplementedError(element_type)
    if np.abs(N - np.eye(npe)).max() > 1e-10:
        raise _lib.FlorenceAMDError("local node coordinates do not match the bases")
    return np.asarray(xi, dtype=np.float64)


def coarse_degrees(element_type, p):
    """Degrees of the coarse levels below a fine degree p."""
    if element_type == "hex":
        return {4: [2, 1], 3: [1], 2: [1]}.get(p, [])
    return [1] if p >= 2 else []


def _exponents(element_type, p):
    if element_type == "hex":
        return [(i, j, k) for k in range(p + 1) for j in range(p + 1) for i in range(p + 1)]
    return [(i, j, k) for s in range(p + 1) for i in range(s + 1) for j in range(s + 1 - i) for k in [s - i - j]]


def _monomials(ex, z):
    """Values and gradients (3, npts, nm) of the monomials x^i y^j z^k at points z."""
    z = np.asarray(z, dtype=np.float64)
    V = np.stack([z[:, 0] ** i * z[:, 1] ** j * z[:, 2] ** k for i, j, k in ex], axis=1)
    pw = lambda x, n: n * x ** (n - 1) if n > 0 else np.zeros_like(x)
    G = np.stack([np.stack([pw(z[:, 0], i) * z[:, 1] ** j * z[:, 2] ** k for i, j, k in ex], axis=1),
                  np.stack([z[:, 0] ** i * pw(z[:, 1], j) * z[:, 2] ** k for i, j, k in ex], axis=1),
                  np.stack([z[:, 0] ** i * z[:, 1] ** j * pw(z[:, 2], k) for i, j, k in ex], axis=1)])
    return V, G


def coarse_basis(element_type, pc, nodes, points):
    """Lagrange basis of degree pc through ``nodes`` (nc, 3): values (npts, nc) and
    gradients (3, npts, nc) at ``points``."""
    ex = _exponents(element_type, pc)
    if len(ex) != nodes.shape[0]:
        raise _lib.FlorenceAMDError("coarse space of degree {} has {} functions for {} nodes".format(
            pc, len(ex), nodes.shape[0]))
    Vn, _ = _monomials(ex, nodes)
    coef = np.linalg.inv(Vn)                     # N_a = sum_m coef[m, a] phi_m
    V, G = _monomials(ex, points)
    return V @ coef, np.einsum("dpm,ma->dpa", G, coef)


def coarse_tensor_basis(pc, cxi, points, JmCt):
    """fl_tensor_basis table (uint8 numpy) of a hex coarse level of degree pc whose local
    nodes sit at ``cxi`` (nc, 3) on the lattice {-1, 0, 1} or {-1, 1}, under the fine
    tensor-product rule ``points``: the 1D Lagrange bases through the lattice at the
    rule's 1D points, the fine g_of_t and the level's local/tensor node maps; None when
    the rule is not a tensor product.  The factorisation is checked against ``JmCt``
    (ngauss, nc, 3), the level's gradients at the points."""
    z = np.asarray(points, dtype=np.float64)
    G = z.shape[0]
    u = [np.unique(np.round(z[:, d], 13)) for d in range(3)]
    Q = len(u[0])
    if any(len(v) != Q for v in u) or Q ** 3 != G:
        return None
    qi = np.stack([np.searchsorted(u[d], np.round(z[:, d], 13)) for d in range(3)], axis=1)
    g_of_t = np.full(G, -1, dtype=np.int32)
    g_of_t[(qi[:, 2] * Q + qi[:, 1]) * Q + qi[:, 0]] = np.arange(G, dtype=np.int32)
    if (g_of_t < 0).any():
        return None
    P1 = pc + 1
    lat = np.linspace(-1.0, 1.0, P1)
    f1d = np.zeros((3, 2, P1, Q))
    for d in range(3):
        x = u[d]
        for a in range(P1):
            others = [b for b in range(P1) if b != a]
            f1d[d, 0, a] = np.prod([(x - lat[b]) / (lat[a] - lat[b]) for b in others], axis=0)
            f1d[d, 1, a] = sum(np.prod([(x - lat[b]) / (lat[a] - lat[b]) for b in others if b != c] + [np.ones(Q)],
                                       axis=0) / (lat[a] - lat[c]) for c in others)
    idx = np.rint((np.asarray(cxi) + 1.0) * 0.5 * pc).astype(np.int64)
    if np.abs(lat[idx] - cxi).max() > 1e-12:
        return None
    nc = idx.shape[0]
    tensor_of_l = ((idx[:, 2] * P1 + idx[:, 1]) * P1 + idx[:, 0]).astype(np.int32)
    if nc != P1 ** 3 or not np.array_equal(np.sort(tensor_of_l), np.arange(nc)):
        return None
    local_of_t = np.empty(nc, dtype=np.int32)
    local_of_t[tensor_of_l] = np.arange(nc, dtype=np.int32)
    for m in range(3):
        rec = np.ones((nc, G))
        for d in range(3):
            rec *= f1d[d, int(d == m)][idx[:, d]][:, qi[:, d]]
        ref = np.asarray(JmCt)[:, :, m].T
        if np.abs(rec - ref).max() > 1e-12 * max(1.0, np.abs(ref).max()):
            return None
    ints = [np.array([P1, Q, 0, 0], dtype=np.int32), g_of_t, local_of_t, tensor_of_l]
    if G % 2:
        ints.append(np.zeros(1, dtype=np.int32))
    return np.concatenate([np.concatenate(ints).view(np.uint8), f1d.ravel().view(np.uint8)])


def _coarse_locals(element_type, pc, xi):
    """Local fine nodes lying on the degree-pc lattice of the reference element."""
    if element_type == "hex":
        if pc not in (1, 2):
            raise NotImplementedError("hexahedral coarse levels are the vertices or the p=2 lattice")
        lat = np.array([-1.0, 1.0]) if pc == 1 else np.array([-1.0, 0.0, 1.0])
        on = np.all(np.min(np.abs(xi[:, :, None] - lat[None, None, :]), axis=2) < 1e-12, axis=1)
    else:
        verts = np.array([[-1, -1, -1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1.0]])
        if pc != 1:
            raise NotImplementedError("tetrahedral coarse levels are the vertices")
        on = np.any(np.all(np.abs(xi[:, None, :] - verts[None, :, :]) < 1e-12, axis=2), axis=1)
    return np.nonzero(on)[0]


def _csr_from_first(elements, weights, coarse_elements, n_fine):
    """Prolongation CSR over fine nodes: fine node f takes the weights of the first
    element containing it, W[local, :] over that element's coarse nodes (the nonzero
    ones, in coarse local order).  Built per local index: every fine node whose first
    element holds it at local position l gets row l of W."""
    flat = elements.ravel()
    _, first = np.unique(flat, return_index=True)
    if first.size != n_fine:
        raise _lib.FlorenceAMDError("a node belongs to no element")
    e, l = np.divmod(first, elements.shape[1])
    keep = np.abs(weights) > 1e-14                # (npe, nc)
    counts = keep.sum(axis=1)[l]
    ptr = np.zeros(n_fine + 1, dtype=np.int64)
    np.cumsum(counts, out=ptr[1:])
    idx = np.empty(int(ptr[-1]), dtype=np.int64)
    w = np.empty(int(ptr[-1]), dtype=np.float64)
    for loc in np.unique(l):
        nodes = np.nonzero(l == loc)[0]
        nz = np.nonzero(keep[loc])[0]
        pos = ptr[nodes][:, None] + np.arange(nz.size)[None, :]
        idx[pos] = coarse_elements[e[nodes]][:, nz]
        w[pos] = weights[loc, nz][None, :]
    return ptr, idx, w


def _transpose(ptr, idx, w, n_coarse):
    """The transpose of a CSR over fine nodes as a CSR over coarse nodes, each coarse
    node's list in ascending fine node order."""
    rows = np.repeat(np.arange(ptr.size - 1), np.diff(ptr))
    order = np.argsort(idx, kind="stable")        # rows ascend within each coarse node
    rptr = np.zeros(n_coarse + 1, dtype=np.int64)
    np.cumsum(np.bincount(idx, minlength=n_coarse), out=rptr[1:])
    return rptr, rows[order], w[order]


class _Level(object):
    """A coarse level: connectivity, node ids in the fine mesh, device pattern and
    transfer tables to the next finer level."""

    def __init__(self, device, degree, elements, fine_ids, JmCt, P, R, tensor_basis=None):
        import torch
        lib = _lib.load()
        i32 = torch.int32
        self.degree = degree
        self.nelem, self.npe = elements.shape
        self.nnode = int(fine_ids.size)
        self.fine_ids = torch.as_tensor(fine_ids.astype(np.int64), device=device)
        self.elements = torch.as_tensor(np.ascontiguousarray(elements, dtype=np.int32), device=device)
        self.JmCt = torch.as_tensor(np.ascontiguousarray(JmCt), device=device)
        self.tensor_basis = None if tensor_basis is None else torch.as_tensor(tensor_basis, device=device)
        as32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=device)
        asf = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)
        self.P_ptr, self.P_idx, self.P_w = as32(P[0]), as32(P[1]), asf(P[2])
        self.R_ptr, self.R_idx, self.R_w = as32(R[0]), as32(R[1]), asf(R[2])
        n, ne, npe = self.nnode, self.nelem, self.npe
        self.node_elem_ptr = torch.empty(n + 1, dtype=i32, device=device)
        self.node_elem = torch.empty(ne * npe, dtype=i32, device=device)
        self.node_nbr_ptr = torch.empty(n + 1, dtype=i32, device=device)
        ws = torch.empty(int(lib.fl_pattern_workspace_size(n)), dtype=torch.uint8, device=device)
        nnz = ctypes.c_int64(0)
        maxdeg = ctypes.c_int32(0)
        P_ = _lib.ptr
        _lib.check(lib.fl_pattern_graph(P_(self.elements), ne, npe, n, P_(self.node_elem_ptr), P_(self.node_elem),
                                        P_(self.node_nbr_ptr), ctypes.byref(nnz), ctypes.byref(maxdeg), P_(ws),
                                        ws.numel(), _lib.stream_ptr()), "fl_pattern_graph (coarse level)")
        del ws
        self.node_nnz = int(nnz.value)
        self.node_nbrs = torch.empty(self.node_nnz, dtype=i32, device=device)
        self.indptr = torch.empty(3 * n + 1, dtype=i32, device=device)
        _lib.check(lib.fl_pattern_fill(P_(self.elements), ne, npe, n, P_(self.node_elem_ptr), P_(self.node_elem),
                                       P_(self.node_nbr_ptr), P_(self.node_nbrs), P_(self.indptr), None,
                                       _lib.stream_ptr()), "fl_pattern_fill (coarse level)")
        nb = _lib.FL_GATHER_BUCKETS
        self.node_order = torch.empty(n, dtype=i32, device=device)
        self.bucket_offsets = (ctypes.c_int32 * (nb + 1))()
        self.bucket_degree = (ctypes.c_int32 * nb)()
        _lib.check(lib.fl_pattern_gather_plan(P_(self.node_nbr_ptr), n, P_(self.node_order), self.bucket_offsets,
                                              self.bucket_degree, _lib.stream_ptr()), "fl_pattern_gather_plan")
        self.status = torch.zeros(2, dtype=i32, device=device)
        self.ranks = torch.empty(int(lib.fl_pattern_rank_table_bytes(ne, npe)) // 2, dtype=torch.int16,
                                 device=device)
        _lib.check(lib.fl_pattern_rank_table(P_(self.elements), ne, npe, P_(self.node_nbr_ptr), P_(self.node_nbrs),
                                             P_(self.ranks), P_(self.status), _lib.stream_ptr()),
                   "fl_pattern_rank_table (coarse level)")
        self.V = torch.empty(9 * self.node_nnz, dtype=torch.float64, device=device)
        self.ws = torch.empty(int(lib.fl_mg_galerkin_workspace_size(ne, npe, n)), dtype=torch.uint8, device=device)
        import torch
        self.dofs = (3 * self.fine_ids[:, None] + torch.arange(3, device=device)[None, :]).reshape(-1)
        self.free = None

    def set_free(self, free_fine):
        """Coarse free dofs: those of the same node in the fine mesh (formed on every
        solve: the caller's mask may change between solves)."""
        self.free = free_fine[self.dofs].contiguous()
        return self.free


def hierarchy_tables(element_type, elements, nnode, quadrature_points):
    """Host tables of the coarse levels below a mesh of ``elements`` (nelem, npe):
    one dict per level, coarsest last, with its degree, connectivity ``elements``
    (coarse node ids), ``fine_ids`` (the fine-mesh node of each coarse node), ``JmCt``
    (ngauss, npe_c, 3) -- the coarse basis gradients at the fine quadrature points --
    ``tensor_basis`` (hexes under a tensor-product rule: the level's fl_tensor_basis
    table, else None) and the prolongation ``P`` = (ptr, idx, w) over the next finer
    level's nodes with its transpose ``R``."""
    el = np.asarray(elements, dtype=np.int64)
    npe = el.shape[1]
    xi = local_node_coordinates(element_type, npe)
    p = int(round(npe ** (1.0 / 3.0))) - 1 if element_type == "hex" else {4: 1, 10: 2, 20: 3, 35: 4}.get(npe, 0)
    degrees = coarse_degrees(element_type, p)
    if not degrees:
        raise _lib.FlorenceAMDError("no coarse space below {} p={}".format(element_type, p))
    z = np.asarray(quadrature_points, dtype=np.float64)
    out = []
    fine_el, fine_xi, fine_n, fine_ids = el, xi, int(nnode), np.arange(int(nnode))
    for pc in degrees:
        loc = _coarse_locals(element_type, pc, fine_xi)
        cxi = fine_xi[loc]
        ids_in_fine, cel = np.unique(fine_el[:, loc], return_inverse=True)
        cel = cel.reshape(fine_el.shape[0], loc.size)
        # the coarse basis at the fine quadrature points (gradients) and at the finer
        # level's nodes (prolongation weights)
        _, G = coarse_basis(element_type, pc, cxi, z)
        W, _ = coarse_basis(element_type, pc, cxi, fine_xi)              # (npe_finer, nc)
        P = _csr_from_first(fine_el, W, cel, fine_n)
        JmCt = np.ascontiguousarray(np.transpose(G, (1, 2, 0)))
        tb = coarse_tensor_basis(pc, cxi, z, JmCt) if element_type == "hex" else None
        out.append({"degree": pc, "elements": cel, "fine_ids": fine_ids[ids_in_fine], "JmCt": JmCt,
                    "tensor_basis": tb, "P": P, "R": _transpose(P[0], P[1], P[2], ids_in_fine.size)})
        fine_el, fine_xi, fine_n, fine_ids = cel, cxi, ids_in_fine.size, fine_ids[ids_in_fine]
    return out


class PMultigridHierarchy(object):
    """The coarse levels of a DeviceMeshData (built once per mesh and function space)."""

    def __init__(self, dm, element_type, function_space):
        z = np.asarray(function_space.quadrature.points, dtype=np.float64)
        if z.shape[0] != dm.ngauss:
            raise _lib.FlorenceAMDError("quadrature points do not match the mesh data")
        tabs = hierarchy_tables(element_type, dm.elements.cpu().numpy(), dm.nnode, z)
        self.levels = [_Level(dm.device, t["degree"], t["elements"], t["fine_ids"], t["JmCt"], t["P"], t["R"],
                              t["tensor_basis"]) for t in tabs]


class MultigridSolver(object):
    """Newton-step linear solver: p-multigrid preconditioned flexible CG (fl_mg_pcg).

    cheb_degree / cheb_range: Chebyshev smoother on [lmax/range, lmax]; coarse_iterations:
    block-Jacobi CG steps on the coarsest level; power_steps: power iterations for lmax on
    a solver's first solve (cold), warm_power_steps on later ones (started from the
    previous solve's vector: the operator is the next Newton step's); fp32_smoother: an
    FP64 CSR's fine-level smoother reads an FP32 copy of it (the outer iteration keeps the
    FP64 operator, so the solve meets rtol on it); matrix_free: where the assembly left C
    point records (tensor-product hexes), the fine-level smoother applies the tangent
    from them element by element (fl_mf_apply: 54 values per point instead of the CSR's
    values) -- the outer iteration still applies the assembled CSR; smoother_records:
    the precision of the copy of those values the smoother reads, made once per solve
    ("bf16": a quarter of the bytes, "fp32", or "fp64": the records themselves).
    ``rtol``, ``maxit``, ``check_every`` and ``iterations`` as LinearSolver."""

    def __init__(self, rtol=1e-12, maxit=2000, check_every=1, cheb_degree=7, cheb_range=60.0,
                 coarse_iterations=60, power_steps=12, warm_power_steps=3, fp32_smoother=True, matrix_free=True,
                 smoother_records="bf16"):
        self.linear_solver = "pmg"
        self.rtol = rtol
        self.maxit = maxit
        self.check_every = check_every
        self.cheb_degree = cheb_degree
        self.cheb_range = cheb_range
        self.coarse_iterations = coarse_iterations
        self.power_steps = power_steps
        self.warm_power_steps = warm_power_steps
        self.fp32_smoother = fp32_smoother
        self.matrix_free = matrix_free
        self.factorised_galerkin = True
        if smoother_records not in ("bf16", "fp32", "fp64"):
            raise ValueError("smoother_records must be bf16, fp32 or fp64")
        self.smoother_records = smoother_records
        self._ye = None
        self._c32 = None
        self.iterations = []
        self._ws = None
        self._h = None
        self._v32 = None
        self._warm_key = None

    def hierarchy(self, dm):
        if self._h is None or self._h[0] is not dm:
            et = dm.element_type
            self._h = (dm, PMultigridHierarchy(dm, et, dm.function_space))
        return self._h[1]

    def galerkin(self, dm, precision="fp64"):
        """Coarse operators of every level from the last assembly's point records."""
        lib = _lib.load()
        h = self.hierarchy(dm)
        ws = dm.workspace()
        rec = int(lib.fl_df_records_form(_lib.PRECISIONS[precision], _lib.ptr(dm.tensor_basis), dm.npe, dm.ngauss))
        P_ = _lib.ptr
        for lv in h.levels:
            tb = lv.tensor_basis if self.factorised_galerkin else None
            _lib.check(lib.fl_mg_galerkin(P_(ws), ws.numel(), dm.ngauss, dm.nelem, dm.npe, rec, P_(lv.JmCt), P_(tb),
                                          lv.npe,
                                          P_(lv.elements), lv.nnode, P_(lv.node_elem_ptr), P_(lv.node_elem),
                                          P_(lv.node_nbr_ptr), P_(lv.node_nbrs), P_(lv.node_order),
                                          lv.bucket_offsets, lv.bucket_degree, P_(lv.indptr), P_(lv.ranks),
                                          P_(lv.V), P_(lv.ws), lv.ws.numel(), P_(lv.status), _lib.stream_ptr()),
                       "fl_mg_galerkin")
        return h

    def Solve(self, dm, V, free_mask, b, x, precision="fp64"):
        import torch
        lib = _lib.load()
        h = self.galerkin(dm, precision)
        nl = 1 + len(h.levels)
        levels = (FlMgLevel * nl)()
        P_ = lambda t: None if t is None else t.data_ptr()
        levels[0] = FlMgLevel(dm.nnode, P_(dm.indptr), P_(dm.node_nbr_ptr), P_(dm.node_nbrs), P_(V), P_(free_mask),
                              None, None, None, None, None, None)
        for i, lv in enumerate(h.levels):
            fr = lv.set_free(free_mask)
            levels[i + 1] = FlMgLevel(lv.nnode, P_(lv.indptr), P_(lv.node_nbr_ptr), P_(lv.node_nbrs), P_(lv.V),
                                      P_(fr), P_(lv.P_ptr), P_(lv.P_idx), P_(lv.P_w), P_(lv.R_ptr), P_(lv.R_idx),
                                      P_(lv.R_w))
        nn = (ctypes.c_int64 * nl)(dm.nnode, *[lv.nnode for lv in h.levels])
        need = int(lib.fl_mg_workspace_size(nn, nl))
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device=dm.device)
        v32 = V.dtype == torch.float32
        mf = None
        if self.matrix_free:
            # the smoother's copy of the C values: BF16 (27 words per point) or FP32 (54)
            c32 = c16 = None
            if self.smoother_records in ("bf16", "fp32"):
                n32 = dm.nelem * (27 if self.smoother_records == "bf16" else 54) * dm.ngauss
                if self._c32 is None or self._c32.numel() < n32 or self._c32.device != dm.device:
                    self._c32 = torch.empty(n32, dtype=torch.float32, device=dm.device)
                if self.smoother_records == "bf16":
                    c16 = self._c32
                else:
                    c32 = self._c32
            mf, self._ye = matrix_free_operator(dm, precision, self._ye, c32, c16)
        # FP32 copies of the values the smoothers read: level 0 (unless matrix-free or
        # FP32 already) and the intermediate levels (the coarsest one runs CG on its own)
        smooth = None
        if self.fp32_smoother:
            srcs = [None if (mf is not None or v32) else V] + [lv.V for lv in h.levels[:-1]] + [None]
            if self._v32 is None or len(self._v32) != nl:
                self._v32 = [None] * nl
            ptrs = (ctypes.c_void_p * nl)()
            for i, src in enumerate(srcs):
                if src is None:
                    continue
                buf = self._v32[i]
                if buf is None or buf.numel() != src.numel() or buf.device != src.device:
                    buf = self._v32[i] = torch.empty(src.numel(), dtype=torch.float32, device=src.device)
                buf.copy_(src)
                ptrs[i] = buf.data_ptr()
            smooth = ptrs
        # the power vectors in the workspace belong to this hierarchy and workspace
        key = (id(h), self._ws.data_ptr(), tuple(nn), bool(self.fp32_smoother), v32, mf is not None,
               self.smoother_records)
        warm = self._warm_key == key
        it = ctypes.c_int32(0)
        rel = ctypes.c_double(0.0)
        _lib.check(lib.fl_mg_pcg(levels, nl, 1 if v32 else 0, smooth,
                                 None if mf is None else ctypes.byref(mf), _lib.ptr(b), _lib.ptr(x),
                                 float(self.rtol), int(self.maxit), int(self.check_every), int(self.cheb_degree),
                                 float(self.cheb_range), int(self.coarse_iterations),
                                 int(self.warm_power_steps if warm else self.power_steps), 1 if warm else 0,
                                 _lib.ptr(self._ws), self._ws.numel(), ctypes.byref(it), ctypes.byref(rel),
                                 _lib.stream_ptr()),
                   "fl_mg_pcg")
        self._warm_key = key
        self.iterations.append(int(it.value))
        if not (rel.value <= self.rtol) and it.value >= self.maxit:
            raise _lib.FlorenceAMDError("multigrid CG did not reach rtol={} in {} iterations (rel={})".format(
                self.rtol, it.value, rel.value))
        return x
