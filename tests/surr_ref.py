"""NumPy reference of the surrogate MLP (test infrastructure): forward pass,
MSE loss and backward pass in float64, optionally rounding to BF16 exactly
where the GPU training kernel feeds a matrix core (weights W2/W3, hidden
activations, dY and dZ2), so its gradient can be compared with the kernel's
to FP32-accumulation accuracy.

Parameter layout: W1[64][4] (r, alpha_a, alpha_b, bias), W2[64][64], b2[64],
W3[6][64], b3[6].
"""
import numpy as np

H, O = 64, 6
SL = {"W1": (0, 256, (64, 4)), "W2": (256, 4352, (64, 64)), "b2": (4352, 4416, (64,)),
      "W3": (4416, 4800, (6, 64)), "b3": (4800, 4806, (6,))}
NPARAMS = 4806


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return {k: p[a:b].reshape(s) for k, (a, b, s) in SL.items()}


def pack(d):
    p = np.zeros(NPARAMS)
    for k, (a, b, s) in SL.items():
        p[a:b] = np.asarray(d[k]).reshape(-1)
    return p


def bf16(a):
    """round to the nearest BF16 (ties to even), returned as float64"""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def act(z):
    return 1.0 - 2.0 / (np.exp(2.0 * z) + 1.0)


def forward(p, x, emulate=False):
    """normalized inputs x[n][3] -> normalized outputs y[n][6] and the activations"""
    q = unpack(p)
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    X = np.concatenate([np.asarray(x, dtype=np.float64), np.ones((len(x), 1))], axis=1)
    h1 = r(act(X @ q["W1"].T))
    h2 = r(act(h1 @ r(q["W2"]).T + q["b2"]))
    y = h2 @ r(q["W3"]).T + q["b3"]
    return y, (X, h1, h2)


def loss_grad(p, x, t, emulate=False):
    q = unpack(p)
    r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
    y, (X, h1, h2) = forward(p, x, emulate)
    n = len(x)
    e = y - np.asarray(t, dtype=np.float64)
    loss = float((e * e).mean())
    dy = 2.0 * e / (n * O)
    dyb = r(dy)
    g = {"W3": dyb.T @ h2, "b3": dy.sum(0)}
    dz2 = (dyb @ r(q["W3"])) * (1.0 - h2 * h2)
    dz2b = r(dz2)
    g["b2"] = dz2.sum(0)
    g["W2"] = dz2b.T @ h1
    dz1 = (dz2b @ r(q["W2"])) * (1.0 - h1 * h1)
    g["W1"] = dz1.T @ X
    return loss, pack(g)
