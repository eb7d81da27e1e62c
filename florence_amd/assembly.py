"""Assemble: the drop-in replacement of the reference's low-level DF assembly.

Python surface kept from the reference:

* ``Assemble(fem_solver, function_space, formulation, mesh, material, Eulerx, Eulerp)
  -> (stiffness, T[:, None], F, mass)`` (Florence/FiniteElements/Assembly/Assembly.py:25-34)
* ``LowLevelAssembly`` / ``_LowLevelAssembly_`` dispatch on the material class name
  and refusal of materials without a low-level dispatcher (Assembly.py:37-95,
  _LowLevelAssembly_.py:45-83).

Everything between the Python call and the returned CSR runs in
``libflorence_amd.so`` on the GPU: the sparsity pattern is built on the device once
per mesh (cached on ``fem_solver``), and every call assembles K and T in HBM.  The
returned ``stiffness`` is a ``DeviceCSR`` whose ``indptr``/``indices`` equal the
reference pattern and whose data stay on the GPU; ``.to_scipy()`` gives the
reference's scipy view.
"""
import time

import numpy as np

from . import _lib
from .materials import MATERIAL_IDS
from .sparse import DeviceCSR

__all__ = ["Assemble", "LowLevelAssembly", "DeviceMeshData", "device_mesh_data", "AssembleResidual"]


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.FlorenceAMDError("florence_amd assembles on the GPU; no HIP device is visible")
    return torch


def tensor_basis_table(function_space, npe):
    """The fl_tensor_basis table of a Lagrange hex under a tensor-product rule, as a
    uint8 numpy array, or None when the element or the rule is not of that form.
    The factorisation is verified against the function space's Jm."""
    from .bases import LagrangeGaussLobatto1D, NodeArrangementHex
    if getattr(function_space, "element_type", None) != "hex":
        return None
    P1 = int(round(npe ** (1.0 / 3.0)))
    if P1 ** 3 != npe:
        return None
    z = np.asarray(function_space.quadrature.points, dtype=np.float64)
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
    C = P1 - 2
    f1d = np.zeros((3, 2, P1, Q))
    for d in range(3):
        for q in range(Q):
            pts = z[qi[:, d] == q, d]
            N, dN = LagrangeGaussLobatto1D(C, float(pts[0]))
            f1d[d, 0, :, q] = N
            f1d[d, 1, :, q] = dN
    perm = np.asarray(NodeArrangementHex(C))
    if perm.shape != (npe,) or not np.array_equal(np.sort(perm), np.arange(npe)):
        return None
    t = [perm % P1, (perm // P1) % P1, perm // (P1 * P1)]
    Jm = np.asarray(function_space.Jm)
    for m in range(3):
        rec = np.ones((npe, G))
        for d in range(3):
            rec *= f1d[d, int(d == m)][t[d]][:, qi[:, d]]
        if np.abs(rec - Jm[m]).max() > 1e-12 * max(1.0, np.abs(Jm[m]).max()):
            return None
    local_of_t = np.empty(npe, dtype=np.int32)
    local_of_t[perm] = np.arange(npe, dtype=np.int32)
    ints = [np.array([P1, Q, 0, 0], dtype=np.int32), g_of_t, local_of_t, np.asarray(perm, dtype=np.int32)]
    if G % 2:
        ints.append(np.zeros(1, dtype=np.int32))
    return np.concatenate([np.concatenate(ints).view(np.uint8), f1d.ravel().view(np.uint8)])


class DeviceMeshData(object):
    """Device-resident, per-(mesh, function space) inputs of the kernels: coordinates,
    int32 connectivity, basis tables, the node graph and the reference CSR pattern."""

    def __init__(self, mesh, function_space, device="cuda", build_indices=True, build_ranks=True,
                 tensor_basis=True, build_colours=True):
        torch = _torch()
        lib = _lib.load()
        self.device = torch.device(device)
        el = np.asarray(mesh.elements)
        if el.max() >= 2 ** 31 - 1:
            raise _lib.FlorenceAMDError("node ids exceed int32")
        self.nelem, self.npe = int(el.shape[0]), int(el.shape[1])
        self.nnode = int(np.asarray(mesh.points).shape[0])
        self.ngauss = int(function_space.AllGauss.shape[0])
        self.element_type = getattr(mesh, "element_type", None)
        self.function_space = function_space
        self.points = torch.as_tensor(np.ascontiguousarray(mesh.points, dtype=np.float64), device=self.device)
        self.elements = torch.as_tensor(np.ascontiguousarray(el, dtype=np.int32), device=self.device)
        self.Jm = torch.as_tensor(np.ascontiguousarray(function_space.Jm, dtype=np.float64), device=self.device)
        self.AllGauss = torch.as_tensor(np.ascontiguousarray(function_space.AllGauss, dtype=np.float64).ravel(),
                                        device=self.device)
        i32 = torch.int32
        self.node_elem_ptr = torch.empty(self.nnode + 1, dtype=i32, device=self.device)
        self.node_elem = torch.empty(self.nelem * self.npe, dtype=i32, device=self.device)
        self.node_nbr_ptr = torch.empty(self.nnode + 1, dtype=i32, device=self.device)
        ws = torch.empty(int(lib.fl_pattern_workspace_size(self.nnode)), dtype=torch.uint8, device=self.device)
        node_nnz = _lib.ctypes.c_int64(0)
        maxdeg = _lib.ctypes.c_int32(0)
        _lib.check(lib.fl_pattern_graph(_lib.ptr(self.elements), self.nelem, self.npe, self.nnode,
                                        _lib.ptr(self.node_elem_ptr), _lib.ptr(self.node_elem),
                                        _lib.ptr(self.node_nbr_ptr), _lib.ctypes.byref(node_nnz),
                                        _lib.ctypes.byref(maxdeg), _lib.ptr(ws), ws.numel(), _lib.stream_ptr()),
                   "fl_pattern_graph")
        del ws
        self.node_nnz = int(node_nnz.value)
        self.max_degree = int(maxdeg.value)
        self.nnz = 9 * self.node_nnz
        self.node_nbrs = torch.empty(self.node_nnz, dtype=i32, device=self.device)
        self.indptr = torch.empty(3 * self.nnode + 1, dtype=i32, device=self.device)
        self.indices = torch.empty(self.nnz, dtype=i32, device=self.device) if build_indices else None
        _lib.check(lib.fl_pattern_fill(_lib.ptr(self.elements), self.nelem, self.npe, self.nnode,
                                       _lib.ptr(self.node_elem_ptr), _lib.ptr(self.node_elem),
                                       _lib.ptr(self.node_nbr_ptr), _lib.ptr(self.node_nbrs),
                                       _lib.ptr(self.indptr), _lib.ptr(self.indices), _lib.stream_ptr()),
                   "fl_pattern_fill")
        nb = _lib.FL_GATHER_BUCKETS
        self.node_order = torch.empty(self.nnode, dtype=i32, device=self.device)
        self.bucket_offsets = (_lib.ctypes.c_int32 * (nb + 1))()
        self.bucket_degree = (_lib.ctypes.c_int32 * nb)()
        _lib.check(lib.fl_pattern_gather_plan(_lib.ptr(self.node_nbr_ptr), self.nnode, _lib.ptr(self.node_order),
                                              self.bucket_offsets, self.bucket_degree, _lib.stream_ptr()),
                   "fl_pattern_gather_plan")
        self.status = torch.zeros(2, dtype=i32, device=self.device)
        # column positions of every (element, row node, column node) entry: built once,
        # they replace the gather's per-call neighbour-list searches
        self.ranks = None
        if build_ranks and self.max_degree <= 32767:
            self.ranks = torch.empty(int(lib.fl_pattern_rank_table_bytes(self.nelem, self.npe)) // 2,
                                     dtype=torch.int16, device=self.device)
            _lib.check(lib.fl_pattern_rank_table(_lib.ptr(self.elements), self.nelem, self.npe,
                                                 _lib.ptr(self.node_nbr_ptr), _lib.ptr(self.node_nbrs),
                                                 _lib.ptr(self.ranks), _lib.ptr(self.status), _lib.stream_ptr()),
                       "fl_pattern_rank_table")
            if int(self.status[0].item()) & _lib.FL_S_PATTERN:
                raise _lib.FlorenceAMDError("inconsistent sparsity pattern (rank table)")
        self._ws = None
        self._ws_res = None
        # sum-factorised hex stiffness: the 1D bases of a tensor-product rule
        self.tensor_basis = None
        tb = tensor_basis_table(function_space, self.npe) if tensor_basis else None
        if tb is not None:
            if tb.size != int(lib.fl_tensor_basis_bytes(int(round(self.npe ** (1 / 3.))),
                                                        int(round(self.ngauss ** (1 / 3.))))):
                raise _lib.FlorenceAMDError("tensor basis table size mismatch")
            self.tensor_basis = torch.as_tensor(tb, device=self.device)
        # the direct scatter of factorised hexes: element colours and the column-block table
        self.ncolour = 0
        self.colour_offsets = None
        self.colour_elems = None
        self.scatter = None
        if build_colours and self.tensor_basis is not None and self.max_degree <= 32767:
            self._build_colour_plan(np.ascontiguousarray(el, dtype=np.int32))

    def _build_colour_plan(self, el32, max_colours=64):
        torch = _torch()
        lib = _lib.load()
        c = _lib.ctypes
        colour = np.empty(self.nelem, dtype=np.int32)
        elems = np.empty(self.nelem, dtype=np.int32)
        offsets = (c.c_int32 * (max_colours + 1))()
        nc = c.c_int32(0)
        _lib.check(lib.fl_pattern_colour(el32.ctypes.data_as(c.c_void_p), self.nelem, self.npe, self.nnode,
                                         max_colours, colour.ctypes.data_as(c.c_void_p),
                                         elems.ctypes.data_as(c.c_void_p), offsets, c.byref(nc)),
                   "fl_pattern_colour")
        self.ncolour = int(nc.value)
        self.colour_offsets = offsets
        self.colour_of_elem = torch.as_tensor(colour, device=self.device)
        self.colour_elems = torch.as_tensor(elems, device=self.device)
        self.scatter = torch.empty(int(lib.fl_pattern_scatter_table_bytes(self.nelem, self.npe)) // 2,
                                   dtype=torch.int16, device=self.device)
        self.status.zero_()
        _lib.check(lib.fl_pattern_scatter_table(_lib.ptr(self.elements), self.nelem, self.npe,
                                                _lib.ptr(self.node_elem_ptr), _lib.ptr(self.node_elem),
                                                _lib.ptr(self.node_nbr_ptr), _lib.ptr(self.node_nbrs),
                                                _lib.ptr(self.colour_of_elem), _lib.ptr(self.tensor_basis),
                                                _lib.ptr(self.scatter), _lib.ptr(self.status), _lib.stream_ptr()),
                   "fl_pattern_scatter_table")
        if int(self.status[0].item()) & _lib.FL_S_PATTERN:
            raise _lib.FlorenceAMDError("inconsistent sparsity pattern or colouring (scatter table)")

    def workspace(self, residual_only=False):
        torch = _torch()
        lib = _lib.load()
        if residual_only:
            if self._ws_res is None:
                n = int(lib.fl_df_residual_workspace_size(self.ngauss, self.nelem, self.npe))
                self._ws_res = torch.empty(n, dtype=torch.uint8, device=self.device)
            return self._ws_res
        # one workspace sized for the FP64 tangent serves every precision
        if self._ws is None:
            n = int(lib.fl_df_workspace_size(self.ngauss, self.nelem, self.npe))
            if n == 0:
                raise _lib.FlorenceAMDError("unsupported element size npe={}".format(self.npe))
            self._ws = torch.empty(n, dtype=torch.uint8, device=self.device)
        return self._ws

    def release_workspace(self):
        self._ws = None
        self._ws_res = None


def device_mesh_data(fem_solver, mesh, function_space):
    """Per-mesh device data cached on the solver object (the reference caches its
    pattern on fem_solver too, FEMSolver.py:1659-1684)."""
    key = (id(mesh), id(function_space), np.asarray(mesh.elements).shape)
    cache = getattr(fem_solver, "_fl_device_cache", None)
    if cache is None or cache[0] != key:
        dm = DeviceMeshData(mesh, function_space)
        setattr(fem_solver, "_fl_device_cache", (key, dm))
        return dm
    return cache[1]


def material_table(dm, material):
    """Device (elem_material, table) of a law with per-element data (the fibre
    directions of AnisotropicMooneyRivlin_1), cached on the mesh data per material
    object and orientation array; (None, None) for the other laws."""
    torch = _torch()
    ori = getattr(material, "anisotropic_orientations", None)
    key = (id(material), id(ori), tuple(material.device_parameters()))
    cache = getattr(dm, "_mat_table", None)
    if cache is not None and cache[0] == key:
        return cache[1], cache[2]
    tab = material.device_table(dm.nelem)
    if tab is None:
        return None, None
    em = torch.as_tensor(tab[0], device=dm.device)
    mt = torch.as_tensor(np.ascontiguousarray(tab[1]), device=dm.device)
    dm._mat_table = (key, em, mt)
    return em, mt


def _as_device_coords(Eulerx, device):
    torch = _torch()
    if isinstance(Eulerx, torch.Tensor):
        x = Eulerx.to(device=device, dtype=torch.float64)
        return x.contiguous()
    return torch.as_tensor(np.ascontiguousarray(Eulerx, dtype=np.float64), device=device)


def assemble_device(dm, material, Eulerx, V=None, T=None, requires_geometry_update=1, stream=None,
                    precision="fp64", csr="fp64", direct=True):
    """Raw device assembly on DeviceMeshData: returns (V, T, status) torch tensors.
    ``precision`` selects the element-block contraction: "fp64" (the reference's
    arithmetic), "fp32" (FP32 MFMA) or "bf16" (BF16 MFMA), FP32 accumulation; T and the
    V sums stay FP64.  ``csr="fp32"`` (with fp32/bf16) stores V as float32.  ``direct``
    (FP64 hexes with a colour plan): the stiffness kernel writes V itself, colour by
    colour, instead of element blocks that a gather sums per node."""
    torch = _torch()
    lib = _lib.load()
    x = _as_device_coords(Eulerx, dm.device)
    vdtype = torch.float32 if csr == "fp32" else torch.float64
    if V is None:
        V = torch.empty(dm.nnz, dtype=vdtype, device=dm.device)
    if V.dtype != vdtype:
        raise ValueError("V dtype {} does not match csr={}".format(V.dtype, csr))
    if T is None:
        T = torch.empty(3 * dm.nnode, dtype=torch.float64, device=dm.device)
    ws = dm.workspace()
    prm, _keep = _lib.double_array(material.device_parameters())
    mid = MATERIAL_IDS.get(type(material).__name__)
    if mid is None:
        raise NotImplementedError("no low-level assembler for material {}".format(type(material).__name__))
    prec = _lib.PRECISIONS[precision] | (_lib.FL_PREC_CSR32 if csr == "fp32" else 0)
    em, mt = material_table(dm, material)
    args = (mid, prm, _lib.ptr(em), _lib.ptr(mt), _lib.ptr(dm.points), _lib.ptr(dm.elements), _lib.ptr(x), _lib.ptr(dm.Jm),
            _lib.ptr(dm.AllGauss), dm.ngauss, dm.nelem, dm.npe, dm.nnode, int(requires_geometry_update),
            _lib.ptr(dm.node_elem_ptr), _lib.ptr(dm.node_elem), _lib.ptr(dm.node_nbr_ptr), _lib.ptr(dm.node_nbrs),
            _lib.ptr(dm.node_order), dm.bucket_offsets, dm.bucket_degree, _lib.ptr(dm.indptr), _lib.ptr(V),
            _lib.ptr(T), prec, _lib.ptr(dm.ranks), _lib.ptr(dm.tensor_basis))
    tail = (_lib.ptr(ws), ws.numel(), _lib.ptr(dm.status), _lib.stream_ptr(stream))
    if direct and dm.scatter is not None:
        _lib.check(lib.fl_df_assemble_cs(*(args + (dm.ncolour, dm.colour_offsets, _lib.ptr(dm.colour_elems),
                                                   _lib.ptr(dm.scatter)) + tail)), "fl_df_assemble_cs")
    else:
        _lib.check(lib.fl_df_assemble_mt(*(args + tail)), "fl_df_assemble_mt")
    return V, T, dm.status


def residual_device(dm, material, Eulerx, T=None, stream=None):
    """Matrix-free internal force on DeviceMeshData (explicit residual-only pass)."""
    torch = _torch()
    lib = _lib.load()
    x = _as_device_coords(Eulerx, dm.device)
    if T is None:
        T = torch.empty(3 * dm.nnode, dtype=torch.float64, device=dm.device)
    ws = dm.workspace(residual_only=True)
    prm, _keep = _lib.double_array(material.device_parameters())
    mid = MATERIAL_IDS.get(type(material).__name__)
    if mid is None:
        raise NotImplementedError("no low-level assembler for material {}".format(type(material).__name__))
    em, mt = material_table(dm, material)
    _lib.check(lib.fl_df_residual_mt(mid, prm, _lib.ptr(em), _lib.ptr(mt), _lib.ptr(dm.points),
                                     _lib.ptr(dm.elements), _lib.ptr(x), _lib.ptr(dm.Jm), _lib.ptr(dm.AllGauss),
                                     dm.ngauss, dm.nelem, dm.npe, dm.nnode, _lib.ptr(dm.node_elem_ptr),
                                     _lib.ptr(dm.node_elem), _lib.ptr(T), _lib.ptr(dm.tensor_basis), _lib.ptr(ws),
                                     ws.numel(), _lib.ptr(dm.status), _lib.stream_ptr(stream)),
               "fl_df_residual_mt")
    return T, dm.status


def LowLevelAssembly(fem_solver, function_space, formulation, mesh, material, Eulerx, Eulerp):
    """Assembly.py:37-95 (static branch): requires a low-level dispatcher."""
    if not getattr(material, "has_low_level_dispatcher", False):
        raise RuntimeError("Cannot dispatch to low level module, since material {} does not support it".format(
            type(material).__name__))
    if getattr(formulation, "fields", "mechanics") != "mechanics":
        raise NotImplementedError("only the displacement formulation is assembled on the GPU")
    if type(material).__name__ not in MATERIAL_IDS:
        raise NotImplementedError("Turning optimise option on for {} material is not supported yet. "
                                  "Consider 'optimise=False' for now".format(type(material).__name__))
    dm = device_mesh_data(fem_solver, mesh, function_space)
    update = int(getattr(fem_solver, "requires_geometry_update", True))
    V, T, status = assemble_device(dm, material, Eulerx, requires_geometry_update=update)
    stiffness = DeviceCSR(dm.indptr, dm.indices, V, (3 * dm.nnode, 3 * dm.nnode))
    fem_solver.last_assembly_status = status
    return stiffness, T[:, None], [], []


def Assemble(fem_solver, function_space, formulation, mesh, material, Eulerx, Eulerp):
    """(stiffness, T[:, None], F, mass) -- Assembly.py:25-34; records fem_solver.assembly_time."""
    t0 = time.time()
    out = LowLevelAssembly(fem_solver, function_space, formulation, mesh, material, Eulerx, Eulerp)
    fem_solver.assembly_time = time.time() - t0
    return out


def AssembleResidual(fem_solver, function_space, formulation, mesh, material, Eulerx, Eulerp):
    """Residual-only (explicit) internal force vector, (3*nnode, 1) on the GPU --
    the mechanics part of AssembleExplicit (Assembly.py:664-716)."""
    dm = device_mesh_data(fem_solver, mesh, function_space)
    T, status = residual_device(dm, material, Eulerx)
    fem_solver.last_assembly_status = status
    return T[:, None]
