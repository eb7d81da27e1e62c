"""GPU parity of the three other low-level laws of the reference's AOT list
(AOT_Assembler.py:8-14: _LinearElastic_, _NearlyIncompressibleMooneyRivlin_,
_AnisotropicMooneyRivlin_1_) against the FP64 oracle, whose laws are pinned by
tests/test_oracle.py (energy Hessians for the two hyperelastic laws, small-strain
Hessian / rotation / linearisation checks for the linear elastic one).

Bar: CSR pattern bit-exact, K and T within 1e-12 relative (infinity norm), on every
data path the assembler has: the dense point and contraction kernels (tets), the
sum-factorised kernels with element blocks and gather (hex p=2 under a Gauss rule) and
the direct colour-ordered scatter (hex p=4, 6^3 Gauss-Lobatto points).
"""
import numpy as np
import pytest

from cases import cube_mesh, deformed, function_space, hex_p4_cube
from florence_amd.materials import AnisotropicMooneyRivlin_1, LinearElastic, NearlyIncompressibleMooneyRivlin
from oracle import df_numpy as O

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def _material(name, nelem, seed=3):
    if name == "LE":
        return LinearElastic(3, mu=1.0, lamb=10.0)
    if name == "NIMR":
        return NearlyIncompressibleMooneyRivlin(3, mu1=0.6, mu2=0.3, lamb=15.0)
    rng = np.random.default_rng(seed)
    ori = rng.standard_normal((nelem, 3))
    ori /= np.linalg.norm(ori, axis=1)[:, None]
    return AnisotropicMooneyRivlin_1(3, mu1=0.5, mu2=0.4, mu3=0.7, lamb=20.0, anisotropic_orientations=ori)


def _meshes():
    out = []
    for et, p, n in (("tet", 2, 2), ("hex", 2, 2)):
        m = cube_mesh(et, p, n, jitter=0.05, seed=p)
        out.append(("{}{}".format(et, p), m, function_space(m, p)))
    m, fs, _ = hex_p4_cube(n=2)
    out.append(("hex4-gll6", m, fs))
    return out


@pytest.mark.parametrize("mat", ["LE", "NIMR", "AMR1"])
@pytest.mark.parametrize("update", [1, 0])
def test_other_laws_match_oracle(cuda, mat, update):
    import torch
    from florence_amd.assembly import DeviceMeshData, assemble_device, residual_device
    for label, m, fs in _meshes():
        material = _material(mat, m.nelem)
        x = deformed(m.points, 0.04, seed=9)
        dm = DeviceMeshData(m, fs)
        V, T, status = assemble_device(dm, material, x, requires_geometry_update=update)
        torch.cuda.synchronize()
        V, T, status = V.cpu().numpy(), T.cpu().numpy(), status.cpu().numpy()
        ip, ix = O.sparsity_pattern(m.elements, m.nnode)
        assert np.array_equal(dm.indptr.cpu().numpy(), ip) and np.array_equal(dm.indices.cpu().numpy(), ix)
        Vr, Tr = O.assemble_csr(m.points, m.elements, x, fs.Jm, fs.AllGauss, material, ip, ix,
                                requires_geometry_update=update)
        ek = np.abs(V - Vr).max() / np.abs(Vr).max()
        et = np.abs(T - Tr).max() / max(np.abs(Tr).max(), 1e-300)
        path = "direct" if dm.scatter is not None else ("factorised" if dm.tensor_basis is not None else "dense")
        print("{} {} update={} ({}): |dK|/|K|={:.3e} |dT|/|T|={:.3e} status={}".format(
            mat, label, update, path, ek, et, status), flush=True)
        assert ek <= RTOL, label
        if update == 1:
            assert et <= RTOL, label
        else:
            assert np.abs(T).max() == 0.0 and np.abs(Tr).max() == 0.0, label
        assert status[0] == 0, label
        if update == 1:
            # residual-only pass (with the material table for the anisotropic law)
            Tres, _ = residual_device(dm, material, x)
            Tro = O.residual_only(m.points, m.elements, x, fs.Jm, fs.AllGauss, material)
            err = np.abs(Tres.cpu().numpy() - Tro).max() / np.abs(Tro).max()
            print("  residual-only |dT|/|T|={:.3e}".format(err), flush=True)
            assert err <= RTOL, label


def test_anisotropic_law_requires_orientations(cuda):
    """The anisotropic law takes its fibre directions per element: the reference-shaped
    entry point without a material table refuses it with FL_E_ARG."""
    import torch
    from florence_amd import _lib
    from florence_amd.assembly import DeviceMeshData
    m = cube_mesh("tet", 2, 1)
    fs = function_space(m, 2)
    mat = _material("AMR1", m.nelem)
    dm = DeviceMeshData(m, fs)
    lib = _lib.load()
    V = torch.empty(dm.nnz, dtype=torch.float64, device="cuda")
    T = torch.empty(3 * dm.nnode, dtype=torch.float64, device="cuda")
    ws = dm.workspace()
    prm, _keep = _lib.double_array(mat.device_parameters())
    rc = lib.fl_df_assemble(mat.material_id, prm, _lib.ptr(dm.points), _lib.ptr(dm.elements), _lib.ptr(dm.points),
                            _lib.ptr(dm.Jm), _lib.ptr(dm.AllGauss), dm.ngauss, dm.nelem, dm.npe, dm.nnode, 1,
                            _lib.ptr(dm.node_elem_ptr), _lib.ptr(dm.node_elem), _lib.ptr(dm.node_nbr_ptr),
                            _lib.ptr(dm.node_nbrs), _lib.ptr(dm.node_order), dm.bucket_offsets, dm.bucket_degree,
                            _lib.ptr(dm.indptr), _lib.ptr(V), _lib.ptr(T), _lib.ptr(ws), ws.numel(),
                            _lib.ptr(dm.status), _lib.stream_ptr())
    assert rc == _lib.FL_E_ARG
    assert b"material table" in lib.fl_last_error()
