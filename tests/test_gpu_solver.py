"""GPU SpMV and block-Jacobi PCG against scipy on the assembled tangent of the p=2 tet cube."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from florence_amd.workloads import clamp_and_pull, cube_tet_p2

pytestmark = pytest.mark.gpu


def _setup():
    import torch
    from florence_amd.assembly import DeviceMeshData, assemble_device
    m, fs, mat = cube_tet_p2(4)
    dm = DeviceMeshData(m, fs)
    X = np.asarray(m.points)
    x = X + 0.02 * np.sin(np.pi * X[:, [1, 2, 0]])
    V, T, _ = assemble_device(dm, mat, x)
    torch.cuda.synchronize()
    K = sp.csr_matrix((V.cpu().numpy(), dm.indices.cpu().numpy(), dm.indptr.cpu().numpy()), shape=(3 * m.nnode,) * 2)
    return m, dm, V, K


def test_spmv_matches_scipy(cuda):
    import torch
    from florence_amd.solver import spmv
    m, dm, V, K = _setup()
    rng = np.random.default_rng(0)
    xv = rng.standard_normal(K.shape[0])
    x = torch.as_tensor(xv, device=cuda)
    y = torch.empty_like(x)
    spmv(dm, V, x, y)
    ref = K @ xv
    bound = 1e-14 * (abs(K) @ np.abs(xv)) + 1e-300   # summation-order bound per row
    err = np.abs(y.cpu().numpy() - ref)
    print("spmv max err {:.3e} max bound {:.3e}".format(err.max(), bound.max()), flush=True)
    assert np.all(err <= 50 * bound)
    free = np.ones(K.shape[0], dtype=np.uint8)
    free[::7] = 0
    spmv(dm, V, x, y, torch.as_tensor(free, device=cuda))
    M = sp.diags(free.astype(float))
    refm = M @ K @ M @ xv + (1 - free) * xv
    bound = 1e-14 * (abs(K) @ np.abs(xv)) + 1e-300
    assert np.all(np.abs(y.cpu().numpy() - refm) <= 50 * bound + 1e-15 * np.abs(xv))


def test_pcg_solves_reduced_system(cuda):
    import torch
    from florence_amd.solver import LinearSolver
    m, dm, V, K = _setup()
    flags = clamp_and_pull(m)
    free = np.isnan(flags.ravel()).astype(np.uint8)
    cin = np.where(free == 1)[0]
    rng = np.random.default_rng(1)
    b = rng.standard_normal(K.shape[0]) * free
    ref = np.zeros_like(b)
    ref[cin] = spla.spsolve(K[cin, :][:, cin].tocsc(), b[cin])
    ls = LinearSolver(rtol=1e-12)
    bx = torch.as_tensor(b, device=cuda)
    x = torch.zeros_like(bx)
    ls.Solve(dm, V, torch.as_tensor(free, device=cuda), bx, x)
    err = np.abs(x.cpu().numpy() - ref).max() / np.abs(ref).max()
    print("pcg iterations", ls.iterations, "rel err", err, flush=True)
    assert err < 1e-9
    assert np.all(x.cpu().numpy()[free == 0] == 0.0)
