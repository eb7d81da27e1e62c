"""florence_amd.Assemble, the reference's signature (Assembly.py:25-34), on the GPU:
returns (stiffness, T[:, None], F, mass) with F = mass = [] on the static path, records
fem_solver.assembly_time, and stiffness.to_scipy() is the reference's scipy CSR
(_LowLevelAssembly_.py:74-78) equal to the oracle's to 1e-12, with the pattern
bit-exact."""
import numpy as np
import pytest
import scipy.sparse as sp

import florence_amd
from cases import deformed
from oracle import df_numpy as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("et,p,mat", [("tet", 2, "NH"), ("hex", 3, "MR")])
def test_assemble_signature_and_values(cuda, et, p, mat):
    m = florence_amd.Mesh()
    m.Cube(n=2, element_type=et)
    m.GetHighOrderMesh(p=p)
    form = florence_amd.DisplacementFormulation(m)
    fs = form.function_spaces[0]
    material = florence_amd.NeoHookean(3, mu=1.0, lamb=10.0) if mat == "NH" else \
        florence_amd.MooneyRivlin(3, mu1=0.5, mu2=0.5, lamb=20.0)
    solver = florence_amd.FEMSolver(optimise=True)
    x = deformed(m.points, 0.04, seed=2)
    out = florence_amd.Assemble(solver, fs, form, m, material, x, np.zeros(m.points.shape[0]))
    assert isinstance(out, tuple) and len(out) == 4
    K, T, F, M = out
    assert F == [] and M == []
    assert isinstance(K, florence_amd.DeviceCSR) and K.shape == (3 * m.nnode, 3 * m.nnode)
    assert tuple(T.shape) == (3 * m.nnode, 1)
    assert solver.assembly_time > 0.0
    Ks = K.to_scipy()
    assert isinstance(Ks, sp.csr_matrix) and Ks.dtype == np.float64 and Ks.indices.dtype == np.int32
    ip, ix = O.sparsity_pattern(m.elements, m.nnode)
    assert np.array_equal(Ks.indptr, ip) and np.array_equal(Ks.indices, ix)
    Vr, Tr = O.assemble_csr(m.points, m.elements, x, fs.Jm, fs.AllGauss, material, ip, ix)
    ek = np.abs(Ks.data - Vr).max() / np.abs(Vr).max()
    et_ = np.abs(T.cpu().numpy()[:, 0] - Tr).max() / np.abs(Tr).max()
    print("Assemble {}{} {}: |dK|/|K|={:.3e} |dT|/|T|={:.3e} assembly_time={:.4f}s".format(
        et, p, mat, ek, et_, solver.assembly_time), flush=True)
    assert ek <= 1e-12 and et_ <= 1e-12
    # the residual-only companion returns the same internal force
    Tres = florence_amd.AssembleResidual(solver, fs, form, m, material, x, None)
    assert tuple(Tres.shape) == (3 * m.nnode, 1)
    assert np.abs(Tres.cpu().numpy()[:, 0] - Tr).max() <= 1e-12 * np.abs(Tr).max()
