"""FP64 NumPy oracle of the learned warm-start network (12 -> 64 -> 64 -> 64 -> 1, SiLU, softplus)
and of its per-constraint inputs (features, match_previous).

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


def match_previous(I, J, I_prev, J_prev, gamma_prev):
    """Previous-step impulse of each row's rod pair (I, J), 0 where the pair had no row last step: the
    warm-start key is the pair of rod indices (rods keep their order between steps)."""
    prev = {(int(a), int(b)): float(g) for a, b, g in zip(I_prev, J_prev, gamma_prev)}
    return np.array([prev.get((int(a), int(b)), 0.0) for a, b in zip(I, J)], np.float64)


def features(I, J, delta0, nrm, posI, posJ, dirs, length, unc, xprev, dt, wide=False):
    """Per-constraint network inputs (16 float32 per row, 12 used), in the expression order of the
    device kernel: gap; normal (3); contact arms along the rods, s_I / (L_I / 2) and s_J / (L_J / 2)
    with s = arm . direction; dt times the relative velocity of the two contact points under U_nc,
    (v_I + w_I x p_I) - (v_J + w_J x p_J) (3); L_I; L_J; previous-step gamma (0 when None).  Lengths
    and directions are the rods' own (SRC/TubuleBind.hpp:54-67: centre +- (L/2) direction).
    wide=True returns the FP64 values before the float32 rounding."""
    I = np.asarray(I, np.int64)
    J = np.asarray(J, np.int64)
    m = I.size
    n = np.asarray(nrm, np.float64).reshape(m, 3)
    pi = np.asarray(posI, np.float64).reshape(m, 3)
    pj = np.asarray(posJ, np.float64).reshape(m, 3)
    t = np.asarray(dirs, np.float64).reshape(-1, 3)
    L = np.asarray(length, np.float64)
    u = np.asarray(unc, np.float64).reshape(-1, 6)
    ti, tj, Li, Lj, ui, uj = t[I], t[J], L[I], L[J], u[I], u[J]
    si = (pi[:, 0] * ti[:, 0] + pi[:, 1] * ti[:, 1]) + pi[:, 2] * ti[:, 2]
    sj = (pj[:, 0] * tj[:, 0] + pj[:, 1] * tj[:, 1]) + pj[:, 2] * tj[:, 2]
    vr = np.empty((m, 3))
    vr[:, 0] = (ui[:, 0] + (ui[:, 4] * pi[:, 2] - ui[:, 5] * pi[:, 1])) - (uj[:, 0] + (uj[:, 4] * pj[:, 2] - uj[:, 5] * pj[:, 1]))
    vr[:, 1] = (ui[:, 1] + (ui[:, 5] * pi[:, 0] - ui[:, 3] * pi[:, 2])) - (uj[:, 1] + (uj[:, 5] * pj[:, 0] - uj[:, 3] * pj[:, 2]))
    vr[:, 2] = (ui[:, 2] + (ui[:, 3] * pi[:, 1] - ui[:, 4] * pi[:, 0])) - (uj[:, 2] + (uj[:, 3] * pj[:, 1] - uj[:, 4] * pj[:, 0]))
    X = np.zeros((m, 16), np.float64)
    X[:, 0] = np.asarray(delta0, np.float64)
    X[:, 1:4] = n
    with np.errstate(divide="ignore", invalid="ignore"):
        X[:, 4] = np.where(Li > 0.0, si / (0.5 * Li), 0.0)
        X[:, 5] = np.where(Lj > 0.0, sj / (0.5 * Lj), 0.0)
    X[:, 6:9] = vr * dt
    X[:, 9] = Li
    X[:, 10] = Lj
    X[:, 11] = 0.0 if xprev is None else np.asarray(xprev, np.float64)
    return X if wide else X.astype(np.float32)


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
