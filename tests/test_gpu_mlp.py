"""Warm-start MLP kernels against the CPU oracle (oracle/mlp.py).

Tolerances: the device accumulates BF16 x BF16 products in FP32 (the oracle sums
them exactly) and evaluates tanhf; an activation that lands on a bf16 rounding
boundary can then round one ulp (2^-8) the other way in the next product.  Forward
outputs and each gradient block are compared norm-wise at 2e-3 relative; the Adam
step (pure FP32 elementwise) at 1e-6.
"""
import numpy as np
import pytest

from oracle import mlp as omlp

pytestmark = pytest.mark.gpu


def _setup(N, nout=3, seed=0):
    import torch
    rng = np.random.Generator(np.random.PCG64(seed))
    P = omlp.init_params(seed)
    P = (P + rng.normal(scale=0.02, size=P.shape)).astype(np.float32)
    X = np.zeros((N, 32), np.float32)
    X[:, :10] = rng.normal(size=(N, 10))
    T = np.zeros((N, 32), np.float32)
    T[:, :nout] = rng.normal(scale=0.3, size=(N, nout))
    d = lambda a: torch.as_tensor(a, device="cuda")
    return P, X, T, d(P), d(X), d(T)


def _fwd(dP, dX):
    import torch
    from florence_amd import _lib
    lib = _lib.load()
    N = dX.shape[0]
    H1 = torch.empty((N, 128), dtype=torch.float32, device="cuda")
    H2 = torch.empty_like(H1)
    Y = torch.empty((N, 32), dtype=torch.float32, device="cuda")
    _lib.check(lib.fl_mlp_forward(_lib.ptr(dP), _lib.ptr(dX), N, _lib.ptr(H1), _lib.ptr(H2), _lib.ptr(Y),
                                  _lib.stream_ptr()), "fl_mlp_forward")
    return H1, H2, Y


def _bwd(dP, dX, acts, dT, nout):
    import torch
    from florence_amd import _lib
    lib = _lib.load()
    N = dX.shape[0]
    G = torch.empty(int(lib.fl_mlp_param_count()), dtype=torch.float32, device="cuda")
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    ws = torch.empty(int(lib.fl_mlp_workspace_size(N)), dtype=torch.uint8, device="cuda")
    H1, H2, Y = acts
    _lib.check(lib.fl_mlp_backward(_lib.ptr(dP), _lib.ptr(dX), _lib.ptr(H1), _lib.ptr(H2), _lib.ptr(Y),
                                   _lib.ptr(dT), N, nout, _lib.ptr(G), _lib.ptr(loss), _lib.ptr(ws), ws.numel(),
                                   _lib.stream_ptr()), "fl_mlp_backward")
    return G, loss


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("N", [32, 4096, 70016])
def test_forward_matches_oracle(cuda, N):
    from florence_amd import _lib
    assert _lib.load().fl_mlp_param_count() == omlp.NPARAM
    P, X, T, dP, dX, dT = _setup(N)
    H1, H2, Y = _fwd(dP, dX)
    rH1, rH2, rY = omlp.forward(P, X)
    for got, ref in ((H1, rH1), (H2, rH2), (Y, rY)):
        assert _rel(got.cpu().numpy(), ref) < 2e-3
    print("N={} forward rel err Y {:.2e}".format(N, _rel(Y.cpu().numpy(), rY)))


@pytest.mark.parametrize("N,nout", [(64, 3), (8192, 3), (70016, 1)])
def test_backward_matches_oracle(cuda, N, nout):
    import torch
    P, X, T, dP, dX, dT = _setup(N, nout, seed=N)
    acts = _fwd(dP, dX)
    G, loss = _bwd(dP, dX, acts, dT, nout)
    rloss, rg, _ = omlp.loss_and_grad(P, X, T, nout)
    assert float(loss.item()) == pytest.approx(rloss, rel=1e-5)
    g = G.cpu().numpy()
    gp, rp = omlp.unpack(g), omlp.unpack(rg)
    for name, _ in omlp.SHAPES:
        e = _rel(gp[name], rp[name])
        print("N={} {} rel err {:.2e}".format(N, name, e))
        assert e < 2e-3, name
    # deterministic: a second call is bitwise equal
    G2, loss2 = _bwd(dP, dX, acts, dT, nout)
    assert torch.equal(G, G2) and torch.equal(loss, loss2)


def test_backward_rejects_bad_sizes(cuda):
    import torch
    from florence_amd import _lib
    lib = _lib.load()
    P, X, T, dP, dX, dT = _setup(40)
    acts = _fwd(dP, dX)
    G = torch.empty(omlp.NPARAM, dtype=torch.float32, device="cuda")
    loss = torch.empty(1, dtype=torch.float32, device="cuda")
    ws = torch.empty(int(lib.fl_mlp_workspace_size(40)), dtype=torch.uint8, device="cuda")
    rc = lib.fl_mlp_backward(_lib.ptr(dP), _lib.ptr(dX), *[_lib.ptr(a) for a in acts], _lib.ptr(dT), 40, 3,
                             _lib.ptr(G), _lib.ptr(loss), _lib.ptr(ws), ws.numel(), _lib.stream_ptr())
    assert rc == -1 and b"multiple of 32" in lib.fl_last_error()


def test_adam_matches_oracle(cuda):
    import torch
    from florence_amd import _lib
    lib = _lib.load()
    rng = np.random.Generator(np.random.PCG64(5))
    n = omlp.NPARAM
    P = rng.normal(size=n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    dP, dm, dv = (torch.as_tensor(a.copy(), device="cuda") for a in (P, m, v))
    Pr, mr, vr = P.astype(np.float64), m.astype(np.float64), v.astype(np.float64)
    for step in range(1, 4):
        G = rng.normal(size=n).astype(np.float32)
        dG = torch.as_tensor(G, device="cuda")
        _lib.check(lib.fl_mlp_adam(_lib.ptr(dP), _lib.ptr(dG), _lib.ptr(dm), _lib.ptr(dv), 1e-3, 0.9, 0.999, 1e-8,
                                   step, _lib.stream_ptr()), "fl_mlp_adam")
        Pr, mr, vr = omlp.adam(Pr, G.astype(np.float64), mr, vr, step, 1e-3)
    assert np.allclose(dP.cpu().numpy(), Pr, rtol=1e-6, atol=1e-7)


def test_training_reduces_loss(cuda):
    """Full-batch Adam through the device kernels fits a smooth map of the inputs."""
    import torch
    from florence_amd import _lib
    lib = _lib.load()
    N = 4096
    rng = np.random.Generator(np.random.PCG64(11))
    X = np.zeros((N, 32), np.float32)
    X[:, :4] = rng.uniform(-1, 1, size=(N, 4))
    T = np.zeros((N, 32), np.float32)
    T[:, 0] = np.sin(2 * X[:, 0]) * X[:, 1]
    T[:, 1] = X[:, 2] * X[:, 3] - 0.5 * X[:, 0] ** 2
    P = omlp.init_params(2)
    dP, dX, dT = (torch.as_tensor(a, device="cuda") for a in (P, X, T))
    m = torch.zeros_like(dP)
    v = torch.zeros_like(dP)
    losses = []
    for step in range(1, 301):
        acts = _fwd(dP, dX)
        G, loss = _bwd(dP, dX, acts, dT, 2)
        _lib.check(lib.fl_mlp_adam(_lib.ptr(dP), _lib.ptr(G), _lib.ptr(m), _lib.ptr(v), 3e-3, 0.9, 0.999, 1e-8,
                                   step, _lib.stream_ptr()), "fl_mlp_adam")
        if step in (1, 300):
            losses.append(float(loss.item()))
    print("loss {:.3e} -> {:.3e}".format(*losses))
    assert losses[1] < 0.05 * losses[0]


def test_warmstart_feature_and_apply_kernels(cuda):
    import torch
    from florence_amd import _lib
    lib = _lib.load()
    rng = np.random.Generator(np.random.PCG64(9))
    nn = 1000
    X, u1, u2 = (rng.normal(size=(nn, 3)) for _ in range(3))
    norm = np.concatenate([rng.normal(size=10), rng.uniform(0.5, 2, size=10)]).astype(np.float32)
    d = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")
    F = torch.empty((nn, 32), dtype=torch.float32, device="cuda")
    dX, du1, du2, dnorm = d(X), d(u1), d(u2), d(norm)     # held until the kernels ran
    _lib.check(lib.fl_warmstart_features(_lib.ptr(dX), _lib.ptr(du1), _lib.ptr(du2), 0.25, nn,
                                         _lib.ptr(dnorm), _lib.ptr(F), _lib.stream_ptr()), "features")
    raw = np.concatenate([X, u1, u1 - u2, np.full((nn, 1), 0.25)], axis=1)
    ref = ((raw - norm[:10].astype(np.float64)) * norm[10:].astype(np.float64)).astype(np.float32)
    Fh = F.cpu().numpy()
    assert np.array_equal(Fh[:, :10], ref) and np.all(Fh[:, 10:] == 0)
    Y = rng.normal(size=(nn, 32)).astype(np.float32)
    on = rng.normal(size=6).astype(np.float32)
    free = (rng.uniform(size=3 * nn) < 0.8).astype(np.uint8)
    ubc = rng.normal(size=3 * nn)
    u = torch.empty(3 * nn, dtype=torch.float64, device="cuda")
    dY, don, dfree, dubc = d(Y), d(on), d(free), d(ubc)
    _lib.check(lib.fl_warmstart_apply(_lib.ptr(du1), _lib.ptr(du2), _lib.ptr(dY), nn, _lib.ptr(don),
                                      _lib.ptr(dfree), _lib.ptr(dubc), _lib.ptr(u), _lib.stream_ptr()), "apply")
    corr = (on[:3].astype(np.float64) * Y[:, :3].astype(np.float64) + on[3:].astype(np.float64)).reshape(-1)
    refu = np.where(free == 1, 2 * u1.reshape(-1) - u2.reshape(-1) + corr, ubc)
    assert np.allclose(u.cpu().numpy(), refu, rtol=1e-14, atol=1e-14)
