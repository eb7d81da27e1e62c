"""fl_df_residual_noalloc (the residual pass with its workspace in the library's fixed
device array, for callers that cannot allocate) against fl_df_residual_mt with a
caller workspace: the same kernels in the same order, so T must be bitwise equal.
Two calls in a row, the second on a smaller mesh (the array keeps the first call's
contents), on the sum-factorised path (hex p=4 with its tensor basis) and on the dense
path (tet p=2 records in the array), and again after an unrelated assembly."""
import numpy as np
import pytest

from cases import deformed

pytestmark = pytest.mark.gpu


def _noalloc(dm, material, x, T):
    from florence_amd import _lib
    from florence_amd.assembly import material_table
    from florence_amd.materials import MATERIAL_IDS
    lib = _lib.load()
    prm, _keep = _lib.double_array(material.device_parameters())
    em, mt = material_table(dm, material)
    rc = lib.fl_df_residual_noalloc(MATERIAL_IDS[type(material).__name__], prm, _lib.ptr(em), _lib.ptr(mt),
                                    _lib.ptr(dm.points), _lib.ptr(dm.elements), _lib.ptr(x), _lib.ptr(dm.Jm),
                                    _lib.ptr(dm.AllGauss), dm.ngauss, dm.nelem, dm.npe, dm.nnode,
                                    _lib.ptr(dm.node_elem_ptr), _lib.ptr(dm.node_elem), _lib.ptr(T),
                                    _lib.ptr(dm.tensor_basis), _lib.ptr(dm.status), _lib.stream_ptr())
    _lib.check(rc, "fl_df_residual_noalloc")
    return T


def _case(builder, n, seed):
    import torch
    from florence_amd.assembly import DeviceMeshData
    mesh, fs, mat = builder(n, seed)
    dm = DeviceMeshData(mesh, fs)
    x = torch.as_tensor(np.ascontiguousarray(deformed(mesh.points, 0.04, seed=seed + 1)), device=dm.device).reshape(-1)
    return dm, mat, x


@pytest.mark.parametrize("form", ["hex_p4_factorised", "tet_p2_dense"])
def test_noalloc_residual_equals_workspace_entry(cuda, form):
    import torch
    from florence_amd.assembly import assemble_device, residual_device
    from florence_amd.workloads import cube_hex_p4, cube_tet_p2
    builder = cube_hex_p4 if form == "hex_p4_factorised" else cube_tet_p2
    sizes = (3, 2) if form == "hex_p4_factorised" else (4, 2)
    for k, n in enumerate(sizes):
        dm, mat, x = _case(builder, n, 10 + k)
        if form == "hex_p4_factorised":
            assert dm.tensor_basis is not None
        else:
            assert dm.tensor_basis is None
        T_ws, st = residual_device(dm, mat, x)
        T_ws = T_ws.clone()
        torch.cuda.synchronize()
        assert int(st[0]) == 0
        T_na = torch.full_like(T_ws, float("nan"))
        _noalloc(dm, mat, x, T_na)
        torch.cuda.synchronize()
        assert int(dm.status[0]) == 0
        assert torch.equal(T_na, T_ws), (form, n)
        # and the full assembly's T on the same state (same pass, rounding-level equal)
        V, T_full, _ = assemble_device(dm, mat, x)
        torch.cuda.synchronize()
        assert (T_na - T_full).abs().max().item() <= 1e-12 * T_full.abs().max().item()
        _noalloc(dm, mat, x, T_na)
        torch.cuda.synchronize()
        assert torch.equal(T_na, T_ws), (form, n, "after an assembly")
