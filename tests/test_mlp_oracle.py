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


def _two_steps(n=800, seed=41):
    """Two consecutive oracle steps of a small periodic 3-D system (the second from the Euler-moved
    state), returning what the feature restatement reads."""
    from alens_amd import systems
    from oracle import oracle as O

    rods, cfg = systems.small_3d(n=n, box=4.0 * (n / 1000.0) ** (1 / 3), seed=seed)
    oc = O.cfg_from_dict(cfg)
    r1 = O.step(rods, oc)
    c, t = O.euler(rods, r1["u"], oc)
    rods2 = dict(rods, center=c, dir=t)
    oc.step = 1
    r2 = O.step(rods2, oc)
    return rods2, cfg, r1, r2


def test_feature_restatement_pins():
    """The restated network inputs are pinned by relations that do not reuse its expressions:
    - the normal component of dt times the relative contact velocity equals b - delta0 = dt (D^T U_nc)_k,
      the kinematically pinned rhs of the rows (a wrong sign, arm or cross product fails);
    - the contact arms lie on the rod axes: p_I = x4 (L_I / 2) t_I, |x4|, |x5| <= 1 (a wrong
      normalisation fails);
    - the previous impulse of a row is that of the same rod pair in the last step (matched here by a
      sorted merge of the pair keys, independent of the restatement's lookup), 0 for new pairs."""
    rods, cfg, r1, r2 = _two_steps()
    rw = r2["rows"]
    I, J = r2["I"], r2["J"]
    m = I.size
    assert m > 100
    xprev = MO.match_previous(I, J, r1["I"], r1["J"], r1["gamma"])
    X = MO.features(I, J, rw["delta0"], rw["nrm"], rw["posI"], rw["posJ"], rods["dir"], rods["length"], r2["unc"],
                    xprev, cfg["dt"], wide=True)
    n = rw["nrm"].reshape(m, 3)
    rate = np.sum(n * X[:, 6:9], axis=1)
    scale = np.max(np.abs(rw["b"] - rw["delta0"]))
    assert np.max(np.abs(rate - (rw["b"] - rw["delta0"]))) <= 1e-12 * scale + 1e-18
    assert np.max(np.abs(X[:, 4])) <= 1.0 + 1e-12 and np.max(np.abs(X[:, 5])) <= 1.0 + 1e-12
    t = np.asarray(rods["dir"]).reshape(-1, 3)
    L = np.asarray(rods["length"])
    arm = X[:, 4:5] * (0.5 * L[I])[:, None] * t[I]
    assert np.max(np.abs(arm - rw["posI"].reshape(m, 3))) <= 1e-12
    arm = X[:, 5:6] * (0.5 * L[J])[:, None] * t[J]
    assert np.max(np.abs(arm - rw["posJ"].reshape(m, 3))) <= 1e-12
    assert np.array_equal(X[:, 0], rw["delta0"]) and np.array_equal(X[:, 9], L[I]) and np.array_equal(X[:, 10], L[J])
    # previous impulses by a sorted merge of keys I * 2^32 + J
    k1 = r1["I"].astype(np.int64) * (1 << 32) + r1["J"].astype(np.int64)
    k2 = I.astype(np.int64) * (1 << 32) + J.astype(np.int64)
    assert np.all(np.diff(k1) > 0) and np.all(np.diff(k2) > 0)  # canonical order: ascending, unique
    pos = np.searchsorted(k1, k2)
    hit = (pos < k1.size) & (k1[np.minimum(pos, k1.size - 1)] == k2)
    want = np.where(hit, r1["gamma"][np.minimum(pos, k1.size - 1)], 0.0)
    assert np.array_equal(X[:, 11], want) and hit.sum() > m // 4 and (~hit).sum() > 0
    assert np.all(X[:, 12:] == 0.0)
    Xf = MO.features(I, J, rw["delta0"], rw["nrm"], rw["posI"], rw["posJ"], rods["dir"], rods["length"], r2["unc"],
                     xprev, cfg["dt"])
    assert Xf.dtype == np.float32 and np.array_equal(Xf, X.astype(np.float32))
