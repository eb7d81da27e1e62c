"""Precision form: tangent contracted on FP32 MFMA with FP32 accumulation.

Checks: the FP32 tangent's error against the FP64 oracle is at FP32 level (stated
bar 1e-5 relative, infinity norm); the residual T, computed in FP64 by the same
kernels, is bitwise the FP64 path's; the inexact-Jacobian Newton on the tet p=2 cube
still reaches the 1e-8 relative residual and its displacement agrees with the FP64
oracle's (tests/golden/c1_newton.npz) to 1e-7 (FP32 tangent) and 1e-6 (BF16 tangent:
linear convergence, the last iterate's residual decides how close it lands)."""
import os

import numpy as np
import pytest

from florence_amd.workloads import cube_hex_p4, cube_tet_p2
from oracle import df_numpy as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "c1_newton.npz")


def _deformed(X, amp=0.04):
    return X + amp * np.sin(np.pi * X[:, [1, 2, 0]])


@pytest.mark.parametrize("case", ["tet_p2", "hex_p4"])
def test_fp32_tangent_error(cuda, case):
    import torch
    from florence_amd.assembly import DeviceMeshData, assemble_device
    m, fs, mat = cube_tet_p2(4) if case == "tet_p2" else cube_hex_p4(2)
    x = _deformed(np.asarray(m.points))
    dm = DeviceMeshData(m, fs)
    V64, T64, _ = assemble_device(dm, mat, x)
    V64, T64 = V64.clone(), T64.clone()
    V32, T32, st = assemble_device(dm, mat, x, precision="fp32")
    V32, T32 = V32.clone(), T32.clone()
    V16, T16, st16 = assemble_device(dm, mat, x, precision="bf16")
    torch.cuda.synchronize()
    indptr, indices = O.sparsity_pattern(m.elements, m.nnode)
    Vr, Tr = O.assemble_csr(m.points, m.elements, x, fs.Jm, fs.AllGauss, mat, indptr, indices)
    e64 = np.abs(V64.cpu().numpy() - Vr).max() / np.abs(Vr).max()
    e32 = np.abs(V32.cpu().numpy() - Vr).max() / np.abs(Vr).max()
    e16 = np.abs(V16.cpu().numpy() - Vr).max() / np.abs(Vr).max()
    print("{}: |K64-K|/|K| {:.2e}  |K32-K|/|K| {:.2e}  |Kbf16-K|/|K| {:.2e}".format(case, e64, e32, e16),
          flush=True)
    assert e64 <= 1e-12
    assert e32 <= 1e-5
    assert e16 <= 5e-3      # BF16 operands (unit roundoff 2^-8 = 3.9e-3), FP32 accumulation
    assert torch.equal(T32, T64) and torch.equal(T16, T64)
    assert int(st[0].item()) == 0


@pytest.mark.parametrize("precision,bar", [("fp32", 1e-7), ("bf16", 1e-6)])
def test_mixed_newton_tet_p2_cube(cuda, precision, bar):
    from florence_amd.boundary_condition import BoundaryCondition
    from florence_amd.solver import FEMSolver, LinearSolver
    from florence_amd.workloads import clamp_and_pull
    ref = np.load(GOLD)
    m, fs, mat = cube_tet_p2(8)

    class Form(object):
        ndim = nvar = 3
        fields = "mechanics"
        function_spaces = (fs, None, None)

    bc = BoundaryCondition()
    bc.SetDirichletCriteria(clamp_and_pull, m)
    solver = FEMSolver(optimise=True, number_of_load_increments=5, newton_raphson_tolerance=1e-8,
                       linear_solver=LinearSolver(rtol=1e-12), tangent_precision=precision)
    res = solver.Solve(formulation=Form(), mesh=m, material=mat, boundary_condition=bc)
    assert not res.newton_raphson_failed_to_converge
    disp = res.TotalDisp[:, :, -1].cpu().numpy()
    err = np.abs(disp - ref["disp"][-1]).max() / np.abs(ref["disp"][-1]).max()
    print("{} tangent newton iterations {} (fp64: {}) rel disp err {:.2e} wall {:.3f}s final residuals {}".format(
        precision, res.newton_iterations, ref["iterations"].tolist(), err, res.wall_time,
        ["{:.1e}".format(v[-1]) for v in res.NRConvergence.values()]), flush=True)
    assert err <= bar


def test_csr32_storage_and_solve(cuda):
    """FL_PREC_CSR32: the FP32-stored values are the FP64 sums of the FP32-tangent path
    rounded once (bitwise), T is unchanged, and block-Jacobi PCG on them solves the same
    system to the FP32 level (solution within 1e-5 of the FP64-valued solve)."""
    import torch
    from florence_amd.assembly import DeviceMeshData, assemble_device
    from florence_amd.solver import LinearSolver, spmv
    m, fs, mat = cube_hex_p4(2)
    x = _deformed(np.asarray(m.points))
    dm = DeviceMeshData(m, fs)
    V64, T64, _ = assemble_device(dm, mat, x, precision="fp32")
    V64, T64 = V64.clone(), T64.clone()
    V32, T32, st = assemble_device(dm, mat, x, precision="fp32", csr="fp32")
    torch.cuda.synchronize()
    assert V32.dtype == torch.float32
    assert torch.equal(V32, V64.to(torch.float32)) and torch.equal(T32, T64)
    assert int(st[0].item()) == 0
    X = np.asarray(m.points)
    free = torch.as_tensor((np.abs(X[:, 0]) > 1e-10).repeat(3).astype(np.uint8), device="cuda")
    b = torch.as_tensor(np.sin(np.arange(3 * m.nnode) * 0.37), device="cuda") * free.to(torch.float64)
    xs = {}
    for name, V in (("v64", V64), ("v32", V32)):
        sol = torch.zeros_like(b)
        LinearSolver(rtol=1e-10).Solve(dm, V, free, b, sol)
        r = torch.empty_like(b)
        spmv(dm, V, sol, r, free)
        assert float(torch.linalg.norm(r - b) / torch.linalg.norm(b)) < 1e-9, name
        xs[name] = sol.cpu().numpy()
    err = np.abs(xs["v32"] - xs["v64"]).max() / np.abs(xs["v64"]).max()
    print("csr32 solve vs csr64 solve: rel diff {:.2e}".format(err))
    assert err < 1e-5
