"""Learned warm start: lockstep batched Newton and what the trained predictor gains.

* Members solved together equal members solved alone (the concatenated operator is
  block diagonal): converged displacements agree to 1e-7 relative.  Both stop at a
  residual <= 1e-8 R0, but not at the same iterate (lockstep members keep iterating
  until the slowest converges, and the PCG tolerance is relative to the whole batch),
  so they agree to the Newton tolerance, not to rounding.
* Training on trajectories the solver records (device kernels end to end) lowers the
  loss and the start residual of held-out members below the extrapolated start's.
* The committed predictor (florence_amd/data, trained by ``bench.py --workload
  learned`` on 2000 trajectories) starts held-out members of its family at less than
  half the extrapolated start's residual, never needs more Newton iterations than the
  extrapolated start it falls back to, and every start mode converges to the same
  displacements (1e-8 relative; 1e-7 with inexact solves).  (On 48 members most
  increments take two iterations either way; the iteration gain shows on the harder
  members, where the extrapolated start needs a third.)  The quadratic and cubic
  extrapolations of the last states (no network) run beside it and reach the same state;
  the "poly" start, per member the order whose FP64 residual is smallest, starts at most
  at the quadratic's residual, and the learned start, which falls back to "poly", at most
  at poly's (1e-3 slack: the modes' converged histories differ at the 1e-9 level).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "florence_amd", "data",
                       "warmstart_twist_hex_p4.npz")


def _family(n=1, nincr=12):
    from florence_amd.learned import TwistStretchFamily
    return TwistStretchFamily(n=n, nincr=nincr)


def test_lockstep_equals_single_members(cuda):
    from florence_amd.learned import LockstepNewton
    fam = _family(nincr=4)
    both = LockstepNewton(fam, [3, 4, 5]).solve("previous", record=True)
    assert not both["failed"]
    m = both["trajectory"].shape[1] // 3          # dofs per member
    for k, seed in ((0, 3), (2, 5)):
        one = LockstepNewton(fam, [seed]).solve("previous", record=True)
        assert not one["failed"]
        a = both["trajectory"][-1][k * m:(k + 1) * m]
        b = one["trajectory"][-1]
        assert np.abs(a - b).max() <= 1e-7 * np.abs(b).max()


def test_training_on_recorded_trajectories(cuda):
    from florence_amd.learned import LockstepNewton, WarmStartMLP, build_dataset, train
    fam = _family()
    gen = LockstepNewton(fam, range(48))
    rec = gen.solve("extrap", record=True)
    assert not rec["failed"]
    net = WarmStartMLP(seed=0)
    F, T = build_dataset(net, fam, [(gen.X, gen.free, rec["trajectory"])])
    hist = train(net, F, T, epochs=40, batch=1024, lr=3e-3)
    print("rows", F.shape[0], "train loss", hist[0], "->", hist[-1])
    assert hist[-1] < 0.2 * hist[0]
    held = LockstepNewton(fam, range(10000, 10016))
    rE = held.solve("extrap")
    rL = held.solve("learned", predictor=net)
    print("start residual extrap {:.3e} learned {:.3e}".format(np.mean(rE["start_residual"]),
                                                               np.mean(rL["start_residual"])))
    assert not rE["failed"] and not rL["failed"]
    assert np.mean(rL["start_residual"]) < np.mean(rE["start_residual"])


def test_committed_predictor_gains(cuda):
    from florence_amd.learned import LockstepNewton, WarmStartMLP
    net = WarmStartMLP.load(WEIGHTS)
    fam = _family(n=2, nincr=30)
    held = LockstepNewton(fam, range(2 * 10 ** 6, 2 * 10 ** 6 + 48))
    modes = ("previous", "extrap", "learned", "extrap2", "extrap3", "poly")
    res = {s: held.solve(s, predictor=net, record=True) for s in modes}
    for s, r in res.items():
        assert not r["failed"], s
        print(s, "lockstep", sum(r["lockstep_iterations"]), "member mean", r["member_iterations"].sum(0).mean(),
              "pcg", sum(r["pcg_iterations"]), "start residual",
              np.mean(r["start_residual"]) if r["start_residual"] else None, "chose learned",
              sum(r["chose_learned"]), "wall {:.2f}s".format(r["wall_s"]))
    ref = res["previous"]["trajectory"][-1]
    for s in modes[1:]:
        assert np.abs(res[s]["trajectory"][-1] - ref).max() <= 1e-8 * np.abs(ref).max()
    sE, sL = np.mean(res["extrap"]["start_residual"]), np.mean(res["learned"]["start_residual"])
    assert sL < 0.5 * sE
    s2, sP = np.mean(res["extrap2"]["start_residual"]), np.mean(res["poly"]["start_residual"])
    assert sP <= 1.001 * s2 and sL <= 1.001 * sP
    assert sum(res["learned"]["lockstep_iterations"]) <= sum(res["extrap"]["lockstep_iterations"])
    assert res["learned"]["member_iterations"].sum() <= res["extrap"]["member_iterations"].sum()
    # the inexact-Newton mode reaches the same state
    ad = LockstepNewton(fam, range(2 * 10 ** 6, 2 * 10 ** 6 + 48), forcing="adaptive")
    ra = ad.solve("learned", predictor=net, record=True)
    assert not ra["failed"]
    assert np.abs(ra["trajectory"][-1] - ref).max() <= 1e-7 * np.abs(ref).max()
    print("adaptive: lockstep", sum(ra["lockstep_iterations"]), "pcg", sum(ra["pcg_iterations"]),
          "wall {:.2f}s".format(ra["wall_s"]))
