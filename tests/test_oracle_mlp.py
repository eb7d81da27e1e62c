"""Pins of the warm-start MLP oracle (oracle/mlp.py) on the CPU.

* bf16 rounding equals torch's float32 -> bfloat16 conversion (round to nearest
  even), ties and specials included;
* the analytic gradient (exact float64 mode) equals central finite differences of the
  loss -- a dropped tanh' factor, a transposed product or a wrong scale fails it;
* the bf16-mode gradient stays within bf16 rounding of the exact one;
* the Adam step equals torch.optim.Adam.
"""
import numpy as np
import pytest

from oracle import mlp


def _problem(N=12, nout=3, seed=1):
    rng = np.random.Generator(np.random.PCG64(seed))
    P = mlp.init_params(seed).astype(np.float64)
    P += rng.normal(scale=0.05, size=P.shape)     # non-zero biases too
    X = np.zeros((N, mlp.IP))
    X[:, :10] = rng.normal(size=(N, 10))
    T = np.zeros((N, mlp.OP))
    T[:, :nout] = rng.normal(scale=0.5, size=(N, nout))
    return P, X, T, nout


def test_bf16_round_matches_torch():
    import torch
    rng = np.random.Generator(np.random.PCG64(0))
    a = np.concatenate([rng.normal(size=4000) * 10.0 ** rng.integers(-30, 30, size=4000),
                        [0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, 3.4e38, 1e-40, 1.00390625, 1.01171875]])
    a = a.astype(np.float32)
    # exact ties: mantissa bits below bf16 equal to 0x8000 (both parities of the kept lsb)
    ties = (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    a = np.concatenate([a, ties, -ties])
    ref = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy()
    got = mlp.bf16_round(a)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert np.isnan(mlp.bf16_round(np.array([np.nan], np.float32)))[0]


def test_exact_gradient_matches_finite_differences():
    P, X, T, nout = _problem()
    loss, g, _ = mlp.loss_and_grad(P, X, T, nout, exact=True)
    assert loss == pytest.approx(mlp.loss_only(P, X, T, nout), rel=1e-14)
    rng = np.random.Generator(np.random.PCG64(7))
    # every bias entry of the first columns plus a random sample of every weight block
    idx, o = [], 0
    for name, shape in mlp.SHAPES:
        n = int(np.prod(shape))
        idx.extend(o + rng.choice(n, size=min(n, 40), replace=False))
        o += n
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (mlp.loss_only(Pp, X, T, nout) - mlp.loss_only(Pm, X, T, nout)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_output_columns_beyond_nout_get_no_gradient():
    P, X, T, nout = _problem(nout=2)
    _, g, _ = mlp.loss_and_grad(P, X, T, nout, exact=True)
    p = mlp.unpack(g)
    assert np.all(p["W3"][:, nout:] == 0) and np.all(p["b3"][nout:] == 0)
    assert np.abs(p["W3"][:, :nout]).max() > 0


def test_bf16_gradient_close_to_exact():
    P, X, T, nout = _problem(N=256)
    _, g64, _ = mlp.loss_and_grad(P, X, T, nout, exact=True)
    _, g16, _ = mlp.loss_and_grad(P.astype(np.float32), X.astype(np.float32), T.astype(np.float32), nout)
    err = np.linalg.norm(g16 - g64) / np.linalg.norm(g64)
    assert 1e-5 < err < 3e-2      # rounding is really applied, and only rounding


def test_adam_matches_torch():
    import torch
    rng = np.random.Generator(np.random.PCG64(3))
    P = rng.normal(size=50)
    m = np.zeros(50)
    v = np.zeros(50)
    tp = torch.tensor(P.copy(), requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    for step in range(1, 6):
        G = rng.normal(size=50)
        P, m, v = mlp.adam(P, G, m, v, step, 1e-2)
        tp.grad = torch.tensor(G)
        opt.step()
        assert np.allclose(P, tp.detach().numpy(), rtol=1e-13, atol=1e-15)
