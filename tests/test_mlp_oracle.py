"""Pins of the network oracle: analytic FP64 gradient against central finite differences, and the
bfloat16 rounding (round to nearest, ties to even)."""
import numpy as np

from oracle import mlp_oracle as MO


def test_bf16_rounding_ties_to_even():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)  # bf16 spacing at 1.0
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -8 + 2 ** -20, -2.5, 3.0e-39], np.float32)
    r = MO.bf16(vals)
    assert r[0] == 1.0
    assert r[1] == one                      # tie -> even (1.0)
    assert r[2] == one + 2 * ulp            # tie -> even (1 + 2 ulp)
    assert r[3] == one + ulp                # above the tie -> up
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-39) < 1e-40


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    P = MO.init_params(3).astype(np.float64)
    X = np.zeros((48, 16), np.float32)
    X[:, :12] = rng.normal(size=(48, 12))
    norm = np.concatenate([rng.normal(size=12) * 0.1, 1.0 + rng.random(12), [1.0]]).astype(np.float32)
    t = rng.random(48) * 2.0
    loss, g = MO.loss_and_grad(P, norm, X, t, bf=False)
    idx = np.concatenate([rng.choice(np.arange(MO.PW1, MO.PB1).reshape(64, 16)[:, :12].ravel(), 20),
                          rng.choice(np.arange(MO.PB1, MO.NPARAM), 60)])
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (MO.loss_and_grad(Pp, norm, X, t, bf=False)[0] - MO.loss_and_grad(Pm, norm, X, t, bf=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])
    # padded input columns carry no gradient
    assert np.all(g[MO.PW1:MO.PB1].reshape(64, 16)[:, 12:] == 0.0)
