"""Device p-refinement with the reference's numbering (fl_mesh_refine through
Mesh.GetHighOrderMesh(numbering="reference")) against the numpy restatement of
HighOrderMeshHex / HighOrderMeshTet_SEMISTABLE (oracle/mesh_ref.py, pinned by
tests/test_oracle_mesh.py): connectivity bit-exact, coordinates within 1e-15; and the
assembled operator on the reference-numbered mesh is the first-appearance mesh's
operator with rows and columns permuted (1e-13)."""
import numpy as np
import pytest
import scipy.sparse as sp

from florence_amd.mesh import Mesh
from oracle.mesh_ref import high_order_mesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("et,p,n", [("hex", 4, 4), ("hex", 2, 3), ("hex", 3, 2), ("tet", 2, 4), ("tet", 3, 2),
                                    ("tet", 4, 1)])
def test_reference_numbering_on_device(cuda, et, p, n):
    m = Mesh()
    m.Cube(n=n, element_type=et)
    X, el = np.asarray(m.points).copy(), np.asarray(m.elements, dtype=np.int64).copy()
    Pr, Er = high_order_mesh(X, el, et, p - 1)
    m.GetHighOrderMesh(p=p, numbering="reference")
    print("{} p={} n={}: {} nodes".format(et, p, n, m.points.shape[0]), flush=True)
    assert np.array_equal(np.asarray(m.elements, dtype=np.int64), Er)
    assert np.abs(np.asarray(m.points) - Pr).max() <= 1e-15


def test_cylinder_reference_numbering(cuda):
    m = Mesh()
    m.Cylinder(nlong=3, nrad=2, element_type="hex")
    X, el = np.asarray(m.points).copy(), np.asarray(m.elements, dtype=np.int64).copy()
    Pr, Er = high_order_mesh(X, el, "hex", 2)
    m.GetHighOrderMesh(p=3, numbering="reference")
    assert np.array_equal(np.asarray(m.elements, dtype=np.int64), Er)
    assert np.abs(np.asarray(m.points) - Pr).max() <= 1e-15


def test_assembly_is_permutation_invariant(cuda):
    import torch
    from cases import deformed, function_space
    from florence_amd.assembly import DeviceMeshData, assemble_device
    from florence_amd.materials import MooneyRivlin
    mat = MooneyRivlin(3, mu1=0.5, mu2=0.5, lamb=20.0)
    ms = {}
    for numbering in ("first", "reference"):
        m = Mesh()
        m.Cube(n=2, element_type="hex")
        m.GetHighOrderMesh(p=4, numbering=numbering)
        ms[numbering] = m
    a, b = ms["first"], ms["reference"]
    perm = np.full(a.nnode, -1)
    perm[np.asarray(a.elements, dtype=np.int64).ravel()] = np.asarray(b.elements, dtype=np.int64).ravel()
    x_a = deformed(a.points, 0.04, seed=1)
    x_b = np.empty_like(x_a)
    x_b[perm] = x_a
    out = []
    for m, x in ((a, x_a), (b, x_b)):
        fs = function_space(m, 4, qtype="lobatto", norder=6)
        dm = DeviceMeshData(m, fs)
        V, T, _ = assemble_device(dm, mat, x)
        torch.cuda.synchronize()
        K = sp.csr_matrix((V.cpu().numpy(), dm.indices.cpu().numpy(), dm.indptr.cpu().numpy()),
                          shape=(3 * m.nnode,) * 2)
        out.append((K, T.cpu().numpy()))
    dof = (3 * perm[:, None] + np.arange(3)).ravel()
    Ka, Ta = out[0]
    Kb, Tb = out[1]
    Kbp = Kb[dof][:, dof]
    assert abs(Kbp - Ka).max() <= 1e-13 * abs(Ka).max()
    assert np.abs(Tb[dof] - Ta).max() <= 1e-13 * np.abs(Ta).max()
