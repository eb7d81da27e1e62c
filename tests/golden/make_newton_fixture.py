"""Golden fixtures of the incremental Newton solves, computed by the CPU oracle
(oracle/newton.py: C restatement of the assembly + scipy sparse direct solve, the
reference's StaticSolver / NewtonRaphson semantics, FEMSolver.py:841-1044):

* newton_tet_p2_cube8.npz -- the tet p=2 cube (unit cube, 8^3 x 6 tets, p=2, 11-point rule,
  Neo-Hookean(1, 10), x=0 clamped, x=1 u_x = 0.2, 5 increments, tolerance 1e-8);
* newton_hex_p4_twist_cube2.npz -- the twisted hex p=4 cube at n=2 (2^3 hexes, p=4, 6^3
  Gauss-Lobatto points, Mooney-Rivlin(0.5, 0.5, 20), x=0 clamped, x=1 twisted 30
  degrees about the x axis, 10 increments, tolerance 1e-8).

Each stores the displacement after every increment (disp), the first Newton iterate of
every increment (first), the iteration counts, the relative and absolute residual
histories and the predictor residual norms, so tests/test_oracle_newton.py can check
the solution against equilibrium, the prescribed values and quadratic convergence.

    python tests/golden/make_newton_fixture.py      # ~1 min
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from florence_amd.workloads import clamp_and_pull, clamp_and_twist, cube_hex_p4, cube_tet_p2  # noqa: E402
from oracle import newton  # noqa: E402


def solve(name, m, fs, mat, flags, nincr):
    rec = {}
    TD, NR, iters = newton.solve_static(m, fs, mat, flags, nincr, tol=1e-8, record=rec)
    width = max(len(v) for v in NR.values())
    hist = np.full((nincr, width), np.nan)
    absh = np.full((nincr, width), np.nan)
    for i in range(nincr):
        v = NR["Increment_{}".format(i)]
        hist[i, :len(v)] = v
        absh[i, :len(v)] = rec["abs_residual"][i]
    np.savez_compressed(os.path.join(HERE, name), disp=np.moveaxis(TD, 2, 0), first=np.array(rec["first_iterate"]),
                        iterations=np.array(iters), nr_history=hist, abs_history=absh,
                        predictor_residual=np.array(rec["predictor_residual"]), points=np.asarray(m.points),
                        flags=np.asarray(flags))
    print(name, "iterations", iters)


def main():
    m, fs, mat = cube_tet_p2(8)
    solve("newton_tet_p2_cube8.npz", m, fs, mat, clamp_and_pull(m), 5)
    m, fs, mat = cube_hex_p4(2)
    solve("newton_hex_p4_twist_cube2.npz", m, fs, mat, clamp_and_twist(m, 30.0), 10)


if __name__ == "__main__":
    main()
