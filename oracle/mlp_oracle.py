"""CPU oracle of the search-surrogate MLP (TEST INFRASTRUCTURE ONLY).

12 -> 64 -> 64 -> 64 -> 1, ReLU hidden layers, linear output, mean squared
error, Adam (Kingma & Ba 2015: beta 0.9 / 0.999, eps 1e-8, bias correction).
Flat parameter layout W1[64][12] b1[64] W2[64][64] b2[64] W3[64][64] b3[64]
W4[64] b4, the layout of tm_mlp_* in include/tsetlin_amd.h.  The analytic
gradient here is pinned by central finite differences in tests/test_surrogate.py;
the GPU kernels are checked against it.
"""
import numpy as np

IN, H = 12, 64
SIZES = [("W1", (H, IN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
         ("W4", (H,)), ("b4", (1,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = p[o:o + k].reshape(shape)
        o += k
    return out


def init(seed):
    """He-normal weights, zero biases."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAMS, np.float32)
    w = unpack(p)
    w["W1"][:] = rng.standard_normal((H, IN)) * np.sqrt(2.0 / IN)
    w["W2"][:] = rng.standard_normal((H, H)) * np.sqrt(2.0 / H)
    w["W3"][:] = rng.standard_normal((H, H)) * np.sqrt(2.0 / H)
    w["W4"][:] = rng.standard_normal(H) * np.sqrt(1.0 / H)
    return p


def forward(p, X):
    w = unpack(np.asarray(p, np.float64))
    X = np.asarray(X, np.float64)
    a1 = np.maximum(X @ w["W1"].T + w["b1"], 0)
    a2 = np.maximum(a1 @ w["W2"].T + w["b2"], 0)
    a3 = np.maximum(a2 @ w["W3"].T + w["b3"], 0)
    out = a3 @ w["W4"] + w["b4"][0]
    return out, (X, a1, a2, a3)


def loss(p, X, Y):
    out, _ = forward(p, X)
    return float(np.mean((out - np.asarray(Y, np.float64)) ** 2))


def gradient(p, X, Y):
    """Analytic gradient of the mean squared error."""
    w = unpack(np.asarray(p, np.float64))
    out, (X, a1, a2, a3) = forward(p, X)
    n = X.shape[0]
    dout = 2.0 * (out - np.asarray(Y, np.float64)) / n
    g = {}
    g["W4"] = dout @ a3
    g["b4"] = np.array([dout.sum()])
    d3 = (a3 > 0) * np.outer(dout, w["W4"])
    g["W3"] = d3.T @ a2
    g["b3"] = d3.sum(0)
    d2 = (a2 > 0) * (d3 @ w["W3"])
    g["W2"] = d2.T @ a1
    g["b2"] = d2.sum(0)
    d1 = (a1 > 0) * (d2 @ w["W2"])
    g["W1"] = d1.T @ X
    g["b1"] = d1.sum(0)
    return np.concatenate([g[name].reshape(-1) for name, _ in SIZES])


def adam(p, X, Y, steps, lr, m=None, v=None, t0=0, b1=0.9, b2=0.999, eps=1e-8):
    p = np.asarray(p, np.float64).copy()
    m = np.zeros_like(p) if m is None else m.copy()
    v = np.zeros_like(p) if v is None else v.copy()
    for t in range(t0 + 1, t0 + steps + 1):
        g = gradient(p, X, Y)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        p -= lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, m, v
