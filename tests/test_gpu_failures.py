"""Failure detection on the device and in the Newton driver, forced from the host (the
kernels always run to completion):

* status word: a deformation that inverts elements sets FL_S_NONPOSITIVE_J with the
  number of quadrature points where det F <= 0 the oracle counts (the reference masks
  them with fabs, _KinematicMeasures_.h:94-99); a NaN coordinate sets FL_S_NONFINITE;
  the residual-only pass reports the same;
* FEMSolver's exits (FEMSolver.py:1010-1039, StaticSolver's truncation :908-918):
  the iteration cap, a NaN residual and a user stop function set
  newton_raphson_failed_to_converge and truncate TotalDisp to the converged increments;
* save_incremental_solution writes every saved increment's displacement
  (FEMSolver.py:1613-1627)."""
import numpy as np
import pytest

from cases import cube_mesh, function_space
from florence_amd.materials import NeoHookean
from florence_amd.workloads import clamp_and_pull, cube_hex_p4

pytestmark = pytest.mark.gpu


def _inverting_state(m):
    """x = X with the centre node of a 2^3 cube moved to (2, 2, 2)."""
    X = np.asarray(m.points)
    x = X.copy()
    x[np.argmin(np.abs(X - 0.5).sum(1))] = [2.0, 2.0, 2.0]
    return x


def test_status_counts_inverted_points(cuda):
    import torch
    from florence_amd import _lib
    from florence_amd.assembly import DeviceMeshData, assemble_device, residual_device
    from oracle import df_numpy as O
    m = cube_mesh("hex", 1, 2)
    fs = function_space(m, 1)
    mat = NeoHookean(3, mu=1.0, lamb=10.0)
    x = _inverting_state(m)
    el = np.asarray(m.elements, dtype=np.int64)
    X = np.asarray(m.points)
    _, F, _ = O.kinematic_measures(fs.Jm, fs.AllGauss, X[el], x[el], 1)
    J = np.linalg.det(F)
    expect = int((J <= 0).sum())
    assert expect == 16 and np.abs(J).min() > 0.1
    dm = DeviceMeshData(m, fs)
    for direct in (True, False):
        _, _, st = assemble_device(dm, mat, x, direct=direct)
        torch.cuda.synchronize()
        s = st.cpu().numpy()
        assert s[0] & _lib.FL_S_NONPOSITIVE_J and not s[0] & _lib.FL_S_NONFINITE and int(s[1]) == expect, s
    _, st = residual_device(dm, mat, x)
    torch.cuda.synchronize()
    s = st.cpu().numpy()
    assert s[0] & _lib.FL_S_NONPOSITIVE_J and int(s[1]) == expect, s
    # a valid state clears it again
    assemble_device(dm, mat, X + 0.01 * X)
    torch.cuda.synchronize()
    assert dm.status.cpu().numpy().tolist() == [0, 0]


def test_status_flags_nonfinite(cuda):
    import torch
    from florence_amd import _lib
    from florence_amd.assembly import DeviceMeshData, assemble_device, residual_device
    m, fs, mat = cube_hex_p4(1)
    x = np.asarray(m.points).copy()
    x[7, 1] = np.nan
    dm = DeviceMeshData(m, fs)
    _, _, st = assemble_device(dm, mat, x)
    torch.cuda.synchronize()
    assert st.cpu().numpy()[0] & _lib.FL_S_NONFINITE
    _, st = residual_device(dm, mat, x)
    torch.cuda.synchronize()
    assert st.cpu().numpy()[0] & _lib.FL_S_NONFINITE


class _Form(object):
    ndim = nvar = 3
    fields = "mechanics"

    def __init__(self, fs):
        self.function_spaces = (fs, None, None)


def _solver_run(m, fs, mat, nincr=3, **kw):
    from florence_amd.boundary_condition import BoundaryCondition
    from florence_amd.solver import FEMSolver, LinearSolver
    bc = BoundaryCondition()
    bc.SetDirichletCriteria(clamp_and_pull, m)
    solver = FEMSolver(optimise=True, number_of_load_increments=nincr, newton_raphson_tolerance=1e-8,
                       linear_solver=LinearSolver(rtol=1e-12), **kw)
    return solver, solver.Solve(formulation=_Form(fs), mesh=m, material=mat, boundary_condition=bc)


def _small():
    m = cube_mesh("tet", 2, 2, jitter=0.05, seed=3)
    return m, function_space(m, 2), NeoHookean(3, mu=1.0, lamb=10.0)


def test_iteration_cap_fails_and_truncates(cuda):
    m, fs, mat = _small()
    solver, res = _solver_run(m, fs, mat, maximum_iteration_for_newton_raphson=1)
    assert res.newton_raphson_failed_to_converge and solver.newton_raphson_failed_to_converge
    assert tuple(res.TotalDisp.shape) == (m.nnode, 3, 0)


def test_nan_residual_fails(cuda):
    def poison(solver, Increment, Eulerx, history, *rest):
        if Increment == 1:
            Eulerx = Eulerx.clone()
            Eulerx[5] = float("nan")
        return Eulerx
    m, fs, mat = _small()
    _, res = _solver_run(m, fs, mat, warm_start=poison)
    assert res.newton_raphson_failed_to_converge
    assert tuple(res.TotalDisp.shape) == (m.nnode, 3, 1)
    assert np.isfinite(res.TotalDisp.cpu().numpy()).all()


def test_user_stop_function(cuda):
    calls = []

    def stop(Increment, Iter, rel, abs_, tol):
        calls.append((Increment, Iter))
        return Increment == 2 and Iter == 1
    m, fs, mat = _small()
    _, res = _solver_run(m, fs, mat, user_defined_stop_func=stop)
    assert res.newton_raphson_failed_to_converge
    assert tuple(res.TotalDisp.shape) == (m.nnode, 3, 2)
    assert calls[-1] == (2, 1)


def test_save_incremental_solution(cuda, tmp_path):
    from scipy.io import loadmat
    m, fs, mat = _small()
    base = str(tmp_path / "incr")
    _, res = _solver_run(m, fs, mat, nincr=4, save_incremental_solution=True, incremental_solution_save_frequency=2,
                         incremental_solution_filename=base)
    assert not res.newton_raphson_failed_to_converge
    disp = res.TotalDisp.cpu().numpy()
    for k in range(4):
        path = "{}_{}.mat".format(base, k)
        if k % 2:
            assert not (tmp_path / "incr_{}.mat".format(k)).exists()
            continue
        assert np.array_equal(loadmat(path)["solution"], disp[:, :, k])
