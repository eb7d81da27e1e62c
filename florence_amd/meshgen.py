"""Device p-refinement with the reference's node numbering (libflorence_amd
fl_mesh_refine; GetHighOrderMesh, Mesh.py:1339-1409, through HighOrderMeshHex /
HighOrderMeshTet_SEMISTABLE).  Used by Mesh.GetHighOrderMesh(numbering="reference")."""
import ctypes

import numpy as np

from . import _lib

__all__ = ["refine_on_device"]


def refine_on_device(points, elements, Neval, decimals=10, device="cuda"):
    """(points, elements) of the refined mesh, as host numpy arrays."""
    import torch
    if not torch.cuda.is_available():
        raise _lib.FlorenceAMDError("the reference-numbered high-order mesh is generated on the GPU; "
                                    "no HIP device is visible")
    lib = _lib.load()
    X = np.ascontiguousarray(points, dtype=np.float64)
    el = np.ascontiguousarray(elements, dtype=np.int64)
    nelem, nv = el.shape
    npe = Neval.shape[1]
    if el.min() < 0 or el.max() >= X.shape[0]:
        raise ValueError("element node ids outside the point list")
    dev = torch.device(device)
    Xd = torch.as_tensor(X, device=dev)
    eld = torch.as_tensor(el.astype(np.int32), device=dev)
    Nd = torch.as_tensor(np.ascontiguousarray(Neval, dtype=np.float64), device=dev)
    cap = X.shape[0] + nelem * (npe - nv)
    P = torch.empty((cap, 3), dtype=torch.float64, device=dev)
    E = torch.empty((nelem, npe), dtype=torch.int32, device=dev)
    ws = torch.empty(int(lib.fl_mesh_refine_workspace_size(X.shape[0], nelem, nv, npe)), dtype=torch.uint8,
                     device=dev)
    nn = ctypes.c_int64(0)
    _lib.check(lib.fl_mesh_refine(_lib.ptr(Xd), X.shape[0], _lib.ptr(eld), nelem, nv, _lib.ptr(Nd), npe,
                                  int(decimals), _lib.ptr(P), _lib.ptr(E), ctypes.byref(nn), _lib.ptr(ws), ws.numel(),
                                  _lib.stream_ptr()), "fl_mesh_refine")
    n = int(nn.value)
    return P[:n].cpu().numpy(), E.cpu().numpy().astype(np.int64)
