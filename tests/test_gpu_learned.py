"""Learned warm start: lockstep batched Newton and what the trained predictor gains.

* Members solved together equal members solved alone (the concatenated operator is
  block diagonal): converged displacements agree to 1e-9 (the Newton tolerance is
  1e-8 on the residual).
* Every start mode converges to the same displacements (1e-9).
* A predictor trained on the device on trajectories the solver records cuts the
  start residual and the Newton iterations of held-out members below those of the
  plain extrapolated start, which it falls back to per member.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _family(nincr=12):
    from florence_amd.learned import TwistStretchFamily
    return TwistStretchFamily(n=1, nincr=nincr)


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
        assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()


def test_learned_warm_start_gains(cuda):
    from florence_amd.learned import LockstepNewton, WarmStartMLP, build_dataset, train
    fam = _family()
    gen = LockstepNewton(fam, range(48))
    rec = gen.solve("extrap", record=True)
    assert not rec["failed"]
    net = WarmStartMLP(seed=0)
    F, T = build_dataset(net, fam, [(gen.X, gen.free, rec["trajectory"])])
    hist = train(net, F, T, epochs=30, batch=2048, lr=3e-3)
    print("train loss", hist[0], "->", hist[-1])
    assert hist[-1] < 0.2 * hist[0]
    held = LockstepNewton(fam, range(10000, 10016))
    res = {s: held.solve(s, predictor=net, record=True) for s in ("previous", "extrap", "learned")}
    for s, r in res.items():
        assert not r["failed"], s
        print(s, "lockstep", sum(r["lockstep_iterations"]), "member", int(r["member_iterations"].sum()),
              "start residual", np.mean(r["start_residual"]) if r["start_residual"] else None,
              "chose learned", sum(r["chose_learned"]))
    ref = res["previous"]["trajectory"][-1]
    for s in ("extrap", "learned"):
        assert np.abs(res[s]["trajectory"][-1] - ref).max() <= 1e-9 * np.abs(ref).max()
    sE, sL = np.array(res["extrap"]["start_residual"]), np.array(res["learned"]["start_residual"])
    assert np.mean(sL) < 0.5 * np.mean(sE)
    assert res["learned"]["member_iterations"].sum() < res["extrap"]["member_iterations"].sum()
    assert sum(res["learned"]["chose_learned"]) > 0
