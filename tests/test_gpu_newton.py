"""GPU incremental Newton (FEMSolver.Solve on the device) against the CPU oracle's
solve of the same workload (tests/golden/c1_newton.npz, written by
tests/golden/make_newton_fixture.py).  Bar: same Newton iteration counts and the
converged displacement within 1e-10 (relative to its max norm)."""
import os

import numpy as np
import pytest

from florence_amd.boundary_condition import BoundaryCondition
from florence_amd.solver import FEMSolver, LinearSolver
from florence_amd.workloads import clamp_and_pull, cube_tet_p2

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden", "c1_newton.npz")


class _Formulation(object):
    ndim = nvar = 3
    fields = "mechanics"

    def __init__(self, fs):
        self.function_spaces = (fs, None, None)


def test_tet_p2_cube_newton_matches_oracle(cuda):
    ref = np.load(GOLD)
    m, fs, mat = cube_tet_p2(8)
    assert np.array_equal(np.asarray(m.points), ref["points"]), "mesh generator changed; regenerate the fixture"
    bc = BoundaryCondition()
    bc.SetDirichletCriteria(clamp_and_pull, m)
    solver = FEMSolver(optimise=True, number_of_load_increments=5, newton_raphson_tolerance=1e-8,
                       linear_solver=LinearSolver(rtol=1e-12))
    res = solver.Solve(formulation=_Formulation(fs), mesh=m, material=mat, boundary_condition=bc)
    disp = res.TotalDisp[:, :, -1].cpu().numpy()
    err = np.abs(disp - ref["disp_final"]).max() / np.abs(ref["disp_final"]).max()
    print("newton iterations", res.newton_iterations, "cg", res.cg_iterations, "wall {:.3f}s".format(res.wall_time),
          "rel err {:.3e}".format(err), flush=True)
    assert list(res.newton_iterations) == ref["iterations"].tolist()
    assert err <= 1e-10
    hist = ref["nr_history"]
    for i in range(5):
        got = res.NRConvergence["Increment_{}".format(i)]
        want = hist[i][~np.isnan(hist[i])]
        assert len(got) == len(want)
        # relative residual histories agree down to the solve's own accuracy
        assert np.allclose(got, want, rtol=1e-3, atol=1e-11)
