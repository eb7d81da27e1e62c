"""Precision form: tangent contracted on FP32 MFMA with FP32 accumulation.

Checks: the FP32 tangent's error against the FP64 oracle is at FP32 level (stated
bar 1e-5 relative, infinity norm); the residual T, computed in FP64 by the same
kernels, is bitwise the FP64 path's; the inexact-Jacobian Newton on the tet p=2 cube
still reaches the 1e-8 relative residual and its displacement agrees with the FP64
oracle's (tests/golden/newton_tet_p2_cube8.npz) to 1e-7 (FP32 tangent) and 1e-6 (BF16 tangent:
linear convergence, the last iterate's residual decides how close it lands)."""
import os

import numpy as np
import pytest

from florence_amd.workloads import cube_hex_p4, cube_tet_p2
from oracle import df_numpy as O

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "newton_tet_p2_cube8.npz")


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


@pytest.mark.parametrize("p", [2, 3, 4])
def test_fp32_factorised_tangent(cuda, p):
    """The FP32 tangent of tensor-product hexes runs the sum-factorised kernel in FP32
    (FP32 MFMA, FP32 accumulation, blocks gathered into the FP64 CSR): within the FP32 bar
    of the FP64 oracle and of the dense FP32 contraction (FL_TUNE_STIFF_SF = 1), with T
    bitwise the FP64 path's, and it leaves the C-tensor records the multigrid reads."""
    import torch
    from cases import cube_mesh, function_space
    from florence_amd import _lib
    from florence_amd.assembly import DeviceMeshData, assemble_device
    from florence_amd.materials import MooneyRivlin
    if p == 4:
        m, fs, mat = cube_hex_p4(2)
    else:
        m = cube_mesh("hex", p, 2, jitter=0.05, seed=4)
        fs = function_space(m, p)
        mat = MooneyRivlin(3, mu1=0.5, mu2=0.5, lamb=20.0)
    x = _deformed(np.asarray(m.points))
    dm = DeviceMeshData(m, fs)
    assert dm.tensor_basis is not None
    lib = _lib.load()
    assert lib.fl_df_records_form(_lib.PRECISIONS["fp32"], _lib.ptr(dm.tensor_basis), dm.npe, dm.ngauss) == \
        lib.fl_df_records_form(_lib.PRECISIONS["fp64"], _lib.ptr(dm.tensor_basis), dm.npe, dm.ngauss)
    V64, T64, _ = assemble_device(dm, mat, x)
    V64, T64 = V64.clone(), T64.clone()
    Vsf, Tsf, st = assemble_device(dm, mat, x, precision="fp32")
    Vsf, Tsf = Vsf.clone(), Tsf.clone()
    assert int(st[0].item()) == 0
    try:
        lib.fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 1)
        Vd, _, _ = assemble_device(dm, mat, x, precision="fp32")
        Vd = Vd.clone()
    finally:
        lib.fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 0)
    torch.cuda.synchronize()
    indptr, indices = O.sparsity_pattern(m.elements, m.nnode)
    Vr, _ = O.assemble_csr(m.points, m.elements, x, fs.Jm, fs.AllGauss, mat, indptr, indices)
    s = np.abs(Vr).max()
    e_sf = np.abs(Vsf.cpu().numpy() - Vr).max() / s
    e_d = np.abs(Vd.cpu().numpy() - Vr).max() / s
    e_sd = np.abs(Vsf.cpu().numpy() - Vd.cpu().numpy()).max() / s
    print("hex p{}: |K32sf-K|/|K| {:.2e}  |K32dense-K|/|K| {:.2e}  |sf-dense| {:.2e}".format(p, e_sf, e_d, e_sd),
          flush=True)
    assert e_sf <= 1e-5 and e_sd <= 1e-5
    assert e_sf > 1e-12          # the contraction really ran in FP32
    assert torch.equal(Tsf, T64)


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
    """FL_PREC_CSR32: on the block path the FP32-stored values are the FP64 sums of the
    FP32-tangent path rounded once (bitwise); on the direct scatter (each element's FP32
    contribution added in colour order) the FP32 CSR is within FP32 rounding of the
    FP64-summed one, whose sums equal the block path's to FP64 rounding; T is unchanged;
    block-Jacobi PCG on the FP32 values solves the same system to the FP32 level
    (solution within 1e-5 of the FP64-valued solve)."""
    import torch
    from florence_amd.assembly import DeviceMeshData, assemble_device
    from florence_amd.solver import LinearSolver, spmv
    m, fs, mat = cube_hex_p4(2)
    x = _deformed(np.asarray(m.points))
    dm = DeviceMeshData(m, fs)
    assert dm.scatter is not None
    Vb64, Tb64, _ = assemble_device(dm, mat, x, precision="fp32", direct=False)
    Vb64, Tb64 = Vb64.clone(), Tb64.clone()
    Vb32, Tb32, _ = assemble_device(dm, mat, x, precision="fp32", csr="fp32", direct=False)
    Vb32 = Vb32.clone()
    V64, T64, _ = assemble_device(dm, mat, x, precision="fp32")
    V64, T64 = V64.clone(), T64.clone()
    V32, T32, st = assemble_device(dm, mat, x, precision="fp32", csr="fp32")
    torch.cuda.synchronize()
    assert V32.dtype == torch.float32 and Vb32.dtype == torch.float32
    assert torch.equal(Vb32, Vb64.to(torch.float32)) and torch.equal(Tb32, Tb64)
    s = float(V64.abs().max())
    assert float((V64 - Vb64).abs().max()) <= 1e-13 * s
    e32 = float((V32.to(torch.float64) - V64).abs().max()) / s
    print("FP32 CSR on the direct scatter: |V32 - V64| / |V64| = {:.2e}".format(e32))
    assert e32 <= 4.8e-7      # at most 8 FP32 roundings of a sum (8 x 2^-24)
    assert torch.equal(T32, T64) and torch.equal(T64, Tb64)
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
