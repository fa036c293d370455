"""CPU restatement of the weight surrogate's arithmetic (TEST INFRASTRUCTURE
ONLY: tests/ load it as the checker of libmuca's MLP kernels; nothing in the
product imports it).

The network is 4 -> 128 -> 128 -> 128 -> 1 with SiLU, parameters packed as
W1[128][4], b1, W2[128][128] (W2[o][k]), b2, W3, b3, w4[128], b4.  The GPU
rounds every weight used in a product, and the activations fed to the two
128 x 128 products, to BF16 (round to nearest even); `bf16=True` applies the
same rounding here, with the arithmetic in float64.  `bf16=False` is the
plain float64 network, which tests pin against central finite differences.
Loss: mean squared error over the batch.
"""
import numpy as np

H, IN = 128, 4
OW1 = 0
OB1 = OW1 + H * IN
OW2 = OB1 + H
OB2 = OW2 + H * H
OW3 = OB2 + H
OB3 = OW3 + H * H
OW4 = OB3 + H
OB4 = OW4 + H
NPARAMS = OB4 + 1


def bf16(a):
    """Round float32 values to the nearest BF16 (ties to even), as float64."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(p):
    p = np.asarray(p, np.float64)
    return dict(W1=p[OW1:OB1].reshape(H, IN), b1=p[OB1:OW2], W2=p[OW2:OB2].reshape(H, H),
                b2=p[OB2:OW3], W3=p[OW3:OB3].reshape(H, H), b3=p[OB3:OW4], w4=p[OW4:OB4],
                b4=p[OB4])


def pack(d):
    return np.concatenate([d["W1"].ravel(), d["b1"], d["W2"].ravel(), d["b2"], d["W3"].ravel(),
                           d["b3"], d["w4"], [d["b4"]]])


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def init_params(seed=0):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases."""
    rng = np.random.default_rng(seed)
    d = dict(W1=rng.uniform(-0.5, 0.5, (H, IN)), b1=np.zeros(H),
             W2=rng.uniform(-1, 1, (H, H)) / np.sqrt(H), b2=np.zeros(H),
             W3=rng.uniform(-1, 1, (H, H)) / np.sqrt(H), b3=np.zeros(H),
             w4=rng.uniform(-1, 1, H) / np.sqrt(H), b4=0.0)
    return pack(d).astype(np.float32)


def forward_backward(p, x, t=None, bf16_round=True, w=None):
    """Returns (y, loss, grad) for inputs x [n][4] and targets t [n] (grad
    and loss are None without targets); w weights each squared error."""
    r = bf16 if bf16_round else (lambda a: np.asarray(a, np.float64))
    d = unpack(p)
    x = np.asarray(x, np.float64)
    W1, W2, W3, w4 = r(d["W1"]), r(d["W2"]), r(d["W3"]), r(d["w4"])
    z1 = d["b1"] + x @ W1.T
    a1 = r(z1 * _sig(z1))
    z2 = a1 @ W2.T + d["b2"]
    a2 = r(z2 * _sig(z2))
    z3 = a2 @ W3.T + d["b3"]
    a3 = z3 * _sig(z3)
    y = a3 @ w4 + d["b4"]
    if t is None:
        return y, None, None
    n = len(y)
    e = y - np.asarray(t, np.float64)
    wt = np.ones(n) if w is None else np.asarray(w, np.float64)
    loss = float(np.sum(wt * e * e) / n)
    dy = 2.0 * wt * e / n

    def dsilu(z):
        s = _sig(z)
        return s * (1.0 + z * (1.0 - s))

    g = {}
    g["b4"] = dy.sum()
    g["w4"] = a3.T @ dy
    dz3 = np.outer(dy, w4) * dsilu(z3)
    g["b3"] = dz3.sum(0)
    dz3r = r(dz3)
    g["W3"] = dz3r.T @ a2
    dz2 = (dz3r @ W3) * dsilu(z2)
    g["b2"] = dz2.sum(0)
    dz2r = r(dz2)
    g["W2"] = dz2r.T @ a1
    dz1 = (dz2r @ W2) * dsilu(z1)
    g["b1"] = dz1.sum(0)
    g["W1"] = dz1.T @ x
    return y, loss, pack(g)


def adam_train(p, x, t, steps, batch, lr, lr_final, b1=0.9, b2=0.999, eps=1e-8, w=None):
    """The training schedule of muca_mlp_train restated (minibatches in order,
    cosine learning rate, Adam), for small checks."""
    p = np.asarray(p, np.float64).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    nb = len(x) // batch
    for s in range(steps):
        off = (s % nb) * batch
        lr_s = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * s / steps))
        _, _, g = forward_backward(p, x[off:off + batch], t[off:off + batch],
                                   w=None if w is None else w[off:off + batch])
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        c1, c2 = 1 - b1 ** (s + 1), 1 - b2 ** (s + 1)
        p = p - lr_s * (m / c1) / (np.sqrt(v / c2) + eps)
    return p
