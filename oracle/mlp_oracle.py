"""FP64 NumPy oracle of the learned warm-start network (12 -> 64 -> 64 -> 64 -> 1, SiLU, softplus).

TEST INFRASTRUCTURE ONLY (imported by tests/ as the checker of libalens.so's MFMA kernels).

Parameter layout (flat float32, 9473 values): W1 64x16 row-major (input columns 12..15 zero), b1 64,
W2 64x64, b2 64, W3 64x64, b3 64, w4 64, b4 1.  norm: mean[12], 1/std[12], gamma scale.

With bf16=True every matrix-product operand is rounded to bfloat16 (round to nearest even) exactly
where the device kernels round them (layer inputs and weights of the three hidden layers in the
forward pass; both operands of every backward product); everything else is FP64.  With bf16=False it
is the plain FP64 network, which the tests pin with central finite differences.
"""
import numpy as np

PW1, PB1, PW2, PB2, PW3, PB3, PW4, PB4 = 0, 1024, 1088, 5184, 5248, 9344, 9408, 9472
NPARAM = 9473
NIN = 12


def bf16(x):
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _r(x, on):
    return bf16(x) if on else np.asarray(x, np.float64)


def unpack(P):
    P = np.asarray(P, np.float64)
    return dict(W1=P[PW1:PB1].reshape(64, 16), b1=P[PB1:PW2], W2=P[PW2:PB2].reshape(64, 64), b2=P[PB2:PW3],
                W3=P[PW3:PB3].reshape(64, 64), b3=P[PB3:PW4], w4=P[PW4:PB4], b4=P[PB4])


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def softplus(o):
    return np.where(o > 20.0, o, np.log1p(np.exp(np.minimum(o, 20.0))))


def inputs(norm, X):
    """Standardised, zero-padded inputs, sample-major (B, 16), computed in float32 like the device."""
    X = np.asarray(X, np.float32).reshape(-1, 16)
    mean = np.asarray(norm[:NIN], np.float32)
    istd = np.asarray(norm[NIN:2 * NIN], np.float32)
    Z = np.zeros((len(X), 16), np.float32)
    Z[:, :NIN] = (X[:, :NIN] - mean) * istd
    return Z.astype(np.float64)


def forward(P, norm, X, bf=True):
    w = unpack(P)
    x = inputs(norm, X)
    z1 = _r(x, bf) @ _r(w["W1"], bf).T + w["b1"]
    h1 = silu(z1)
    z2 = _r(h1, bf) @ _r(w["W2"], bf).T + w["b2"]
    h2 = silu(z2)
    z3 = _r(h2, bf) @ _r(w["W3"], bf).T + w["b3"]
    h3 = silu(z3)
    o = h3 @ w["w4"] + w["b4"]
    return dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, h3=h3, o=o, y=softplus(o))


def predict(P, norm, X, bf=True):
    """gamma0 = gamma_scale * network(X)."""
    return float(norm[2 * NIN]) * forward(P, norm, X, bf)["y"]


def loss_and_grad(P, norm, X, target, bf=True):
    """loss = mean (y - t)^2 and its gradient with respect to the flat parameter vector."""
    w = unpack(P)
    f = forward(P, norm, X, bf)
    t = np.asarray(target, np.float64)
    B = len(t)
    d = f["y"] - t
    loss = float(np.mean(d * d))
    dO = 2.0 * d / B / (1.0 + np.exp(-f["o"]))
    g = np.zeros(NPARAM)
    g[PW4:PB4] = f["h3"].T @ dO
    g[PB4] = dO.sum()
    dZ3 = np.outer(dO, w["w4"]) * silu_grad(f["z3"])
    g[PW3:PB3] = (_r(dZ3, bf).T @ _r(f["h2"], bf)).reshape(-1)
    g[PB3:PW4] = dZ3.sum(0)
    dZ2 = (_r(dZ3, bf) @ _r(w["W3"], bf)) * silu_grad(f["z2"])
    g[PW2:PB2] = (_r(dZ2, bf).T @ _r(f["h1"], bf)).reshape(-1)
    g[PB2:PW3] = dZ2.sum(0)
    dZ1 = (_r(dZ2, bf) @ _r(w["W2"], bf)) * silu_grad(f["z1"])
    g1 = _r(dZ1, bf).T @ _r(f["x"], bf)
    g1[:, NIN:] = 0.0
    g[PW1:PB1] = g1.reshape(-1)
    g[PB1:PW2] = dZ1.sum(0)
    return loss, g


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    P = np.zeros(NPARAM, np.float32)

    def glorot(rows, cols, fan_in):
        a = np.sqrt(6.0 / (fan_in + rows))
        W = rng.uniform(-a, a, (rows, cols))
        W[:, fan_in:] = 0.0
        return W

    P[PW1:PB1] = glorot(64, 16, NIN).reshape(-1)
    P[PW2:PB2] = glorot(64, 64, 64).reshape(-1)
    P[PW3:PB3] = glorot(64, 64, 64).reshape(-1)
    P[PW4:PB4] = glorot(1, 64, 64).reshape(-1)
    P[PB1:PW2] = rng.normal(size=64) * 0.1
    P[PB2:PW3] = rng.normal(size=64) * 0.1
    P[PB3:PW4] = rng.normal(size=64) * 0.1
    P[PB4] = 0.5413248546
    return P
