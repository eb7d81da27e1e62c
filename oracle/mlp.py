"""CPU oracle of the warm-start MLP -- TEST INFRASTRUCTURE (see oracle/__init__.py).

The network has no counterpart in the reference; this module restates, in numpy,
what the device kernels compute, so the tests can check the device forward pass,
the backward pass (gradients of the MSE loss) and the Adam update:

    H1 = tanh(bf(X) bf(W1) + b1)      X: (N, 32)  W1: (32, 128)
    H2 = tanh(bf(H1) bf(W2) + b2)     W2: (128, 128)
    Y  = bf(H2) bf(W3) + b3           W3: (128, 32)
    L  = sum_{n, c < nout} (Y - T)^2 / (nout N)

``bf`` rounds a float32 to bfloat16 (round to nearest, ties to even), applied to both
operands of every product, as the BF16 MFMA kernels do; products are accumulated
exactly here (float64 sums of bf16 x bf16 products, which are exact in float64),
where the device accumulates in float32.  With ``exact=True`` no rounding is applied
and everything is float64: that mode is what the finite-difference checks pin.

Backward (straight-through rounding, the gradient of the rounded forward pass):
    dY  = 2 (Y - T) / (nout N) on the first nout columns
    dW3 = bf(H2)^T bf(dY),  db3 = sum_n dY,  dH2 = bf(dY) bf(W3)^T,  dZ2 = dH2 (1 - H2^2)
    dW2 = bf(H1)^T bf(dZ2), db2 = sum_n dZ2, dH1 = bf(dZ2) bf(W2)^T, dZ1 = dH1 (1 - H1^2)
    dW1 = bf(X)^T bf(dZ1),  db1 = sum_n dZ1
"""
import numpy as np

IP, H, OP = 32, 128, 32
SHAPES = [("W1", (IP, H)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, OP)), ("b3", (OP,))]
NPARAM = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(P):
    """Views of the flat parameter vector in the device's order W1 b1 W2 b2 W3 b3."""
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = P[o:o + n].reshape(shape)
        o += n
    return out


def bf16_round(a):
    """float32 -> nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) & 0xFFFF0000).astype(np.uint32)
    nan = np.isnan(a)
    out = r.view(np.float32).copy()
    out[nan] = np.nan
    return out


def _mm(A, B, exact):
    if exact:
        return np.asarray(A, np.float64) @ np.asarray(B, np.float64)
    return bf16_round(A).astype(np.float64) @ bf16_round(B).astype(np.float64)


def _cast(a, exact):
    return np.asarray(a, np.float64) if exact else np.asarray(a, np.float64).astype(np.float32)


def forward(P, X, exact=False):
    p = unpack(np.asarray(P, np.float64 if exact else np.float32))
    H1 = _cast(np.tanh(_mm(X, p["W1"], exact) + p["b1"]), exact)
    H2 = _cast(np.tanh(_mm(H1, p["W2"], exact) + p["b2"]), exact)
    Y = _cast(_mm(H2, p["W3"], exact) + p["b3"], exact)
    return H1, H2, Y


def loss_and_grad(P, X, T, nout, exact=False):
    """(loss, flat gradient, (H1, H2, Y)) of the MSE on the first ``nout`` outputs."""
    p = unpack(np.asarray(P, np.float64 if exact else np.float32))
    H1, H2, Y = forward(P, X, exact)
    N = X.shape[0]
    D = np.zeros_like(Y, dtype=np.float64)
    D[:, :nout] = np.asarray(Y, np.float64)[:, :nout] - np.asarray(T, np.float64)[:, :nout]
    loss = float((D ** 2).sum() / (nout * N))
    dY = _cast(2.0 * D / (nout * N), exact)
    g = {}
    g["W3"] = _mm(H2.T, dY, exact)
    g["b3"] = dY.sum(0, dtype=np.float64)
    dZ2 = _cast(_mm(dY, p["W3"].T, exact) * (1.0 - np.asarray(H2, np.float64) ** 2), exact)
    g["W2"] = _mm(H1.T, dZ2, exact)
    g["b2"] = dZ2.sum(0, dtype=np.float64)
    dZ1 = _cast(_mm(dZ2, p["W2"].T, exact) * (1.0 - np.asarray(H1, np.float64) ** 2), exact)
    g["W1"] = _mm(np.asarray(X).T, dZ1, exact)
    g["b1"] = dZ1.sum(0, dtype=np.float64)
    flat = np.concatenate([np.asarray(g[name], np.float64).reshape(-1) for name, _ in SHAPES])
    return loss, flat, (H1, H2, Y)


def loss_only(P, X, T, nout, exact=True):
    Y = forward(P, X, exact)[2]
    D = np.asarray(Y, np.float64)[:, :nout] - np.asarray(T, np.float64)[:, :nout]
    return float((D ** 2).sum() / (nout * X.shape[0]))


def adam(P, G, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, Algorithm 1, bias-corrected moments, eps added
    to the square root); returns new (P, m, v)."""
    m = beta1 * m + (1.0 - beta1) * G
    v = beta2 * v + (1.0 - beta2) * G * G
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    return P - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed=0):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases (float32)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    parts = []
    for name, shape in SHAPES:
        if name.startswith("W"):
            b = 1.0 / np.sqrt(shape[0])
            parts.append(rng.uniform(-b, b, size=shape).reshape(-1))
        else:
            parts.append(np.zeros(shape).reshape(-1))
    return np.concatenate(parts).astype(np.float32)
