"""Golden fixture of the incremental Newton solve of the tet p=2 cube workload
(unit cube, 8^3 x 6 tets, p=2, 11-point rule, Neo-Hookean(1, 10), x=0 clamped,
x=1 u_x = 0.2, 5 increments, tolerance 1e-8) computed by the CPU oracle
(oracle/newton.py: C restatement of the assembly + scipy sparse direct solve).

    python tests/golden/make_newton_fixture.py      # writes c1_newton.npz (~1 min)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from florence_amd.workloads import clamp_and_pull, cube_tet_p2  # noqa: E402
from oracle import newton  # noqa: E402


def main():
    m, fs, mat = cube_tet_p2(8)
    TD, NR, iters = newton.solve_static(m, fs, mat, clamp_and_pull(m), 5, tol=1e-8)
    hist = np.full((5, max(len(v) for v in NR.values())), np.nan)
    for i in range(5):
        v = NR["Increment_{}".format(i)]
        hist[i, :len(v)] = v
    np.savez_compressed(os.path.join(HERE, "c1_newton.npz"), disp_final=TD[:, :, -1], disp_inc0=TD[:, :, 0],
                        iterations=np.array(iters), nr_history=hist, points=np.asarray(m.points))
    print("iterations", iters)


if __name__ == "__main__":
    main()
