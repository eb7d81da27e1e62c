"""The polynomial starts of the incremental Newton loop (florence_amd.learned): the order-q
extrapolation through q + 1 equally spaced states reproduces every polynomial of degree
<= q exactly at the next load factor and misses degree q + 1 (so a wrong coefficient or
a shifted state fails); the start of increment k uses order min(q, k - 1), the states
it has (u_0 = 0 at load 0 included)."""
import numpy as np
import pytest

from florence_amd.learned import EXTRAP_COEF, _base, extrapolate


@pytest.mark.parametrize("order", [1, 2, 3])
def test_exact_for_polynomials_up_to_its_order(order):
    rng = np.random.default_rng(order)
    k = 7
    for deg in range(order + 2):
        c = rng.normal(size=deg + 1)
        f = lambda t: np.polyval(c, t)
        hist = [np.array([f(k - 1 - j)]) for j in range(order + 1)]
        got = extrapolate(hist, order)[0]
        if deg <= order:
            assert abs(got - f(k)) <= 1e-12 * max(1.0, abs(f(k)))
        else:
            assert abs(got - f(k)) > 1e-6 * max(1.0, abs(f(k)))
    assert len(EXTRAP_COEF[order]) == order + 1 and abs(sum(EXTRAP_COEF[order]) - 1.0) < 1e-15


def test_base_uses_the_states_available():
    torch = pytest.importorskip("torch")
    U = torch.tensor([[float(t) ** 3 - 2.0 * t] for t in range(8)], dtype=torch.float64)
    for k in range(2, 8):
        q = min(3, k - 1)
        want = extrapolate([U[k - 1 - j] for j in range(q + 1)], q)
        assert torch.equal(_base(U, k, 3), want)
        if k >= 4:
            assert float(_base(U, k, 3)[0]) == float(U[k][0])     # a cubic path: exact
    assert float(_base(U, 2, 3)[0]) == 2 * float(U[1][0]) - float(U[0][0])
