"""float64 numpy restatement of the selector MLP (12 -> 32 -> 32 -> 4, ReLU,
ReLU, linear) with its mean-squared-error loss and analytic gradient.
TEST INFRASTRUCTURE ONLY: the checker for the library's HIP forward/backward
kernels; pinned itself by central finite differences in the tests.

Parameter layout: W1[32][12], b1[32], W2[32][32], b2[32], W3[4][32], b3[4].
"""
import numpy as np

NI, H1, H2, NO = 12, 32, 32, 4
SIZES = [("W1", (H1, NI)), ("b1", (H1,)), ("W2", (H2, H1)), ("b2", (H2,)), ("W3", (NO, H2)),
         ("b3", (NO,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, at = {}, 0
    for name, shape in SIZES:
        k = int(np.prod(shape))
        out[name] = p[at:at + k].reshape(shape)
        at += k
    return out


def forward(p, x):
    w = unpack(p)
    x = np.asarray(x, np.float64)
    a1 = x @ w["W1"].T + w["b1"]
    h1 = np.maximum(a1, 0.0)
    a2 = h1 @ w["W2"].T + w["b2"]
    h2 = np.maximum(a2, 0.0)
    y = h2 @ w["W3"].T + w["b3"]
    return y, (x, a1, h1, a2, h2)


def loss(p, x, t):
    y, _ = forward(p, x)
    return float(np.mean((y - np.asarray(t, np.float64)) ** 2))


def loss_grad(p, x, t):
    w = unpack(p)
    y, (x, a1, h1, a2, h2) = forward(p, x)
    t = np.asarray(t, np.float64)
    ns = x.shape[0]
    dy = 2.0 * (y - t) / (ns * NO)
    g = {"W3": dy.T @ h2, "b3": dy.sum(0)}
    dh2 = (dy @ w["W3"]) * (a2 > 0)
    g["W2"] = dh2.T @ h1
    g["b2"] = dh2.sum(0)
    dh1 = (dh2 @ w["W2"]) * (a1 > 0)
    g["W1"] = dh1.T @ x
    g["b1"] = dh1.sum(0)
    flat = np.concatenate([g[name].ravel() for name, _ in SIZES])
    return float(np.mean((y - t) ** 2)), flat
