"""GPU parity of caller-supplied constraint blocks (the reference's constraint pool,
SRC/TubuleSystem.cpp:694-745) and of the constraint stress sums, against the FP64 oracle.

Caller rows are appended after the collision rows; bilateral rows are unbounded springs (1/kappa on
the diagonal) or rigid links; the solve starts from each block's gamma.  Bar: bit-for-bit.
"""
import numpy as np
import pytest

from alens_amd import capi, systems
from oracle import oracle as O

from .test_gpu_parity import assert_bits

pytestmark = pytest.mark.gpu


def make_ext(rods, me, seed, free_len=0.05, kappas=(0.0, 5.0, 50.0)):
    """Springs between random rod pairs: anchors at random points along each rod, the normal along
    the anchor separation, delta0 = |P - Q| - free length, gamma0 = -kappa * delta0 (the reference's
    protein blocks, SRC/TubuleSystem.cpp:713-736); the last row is unilateral with gamma_lb < 0."""
    rng = np.random.default_rng(seed)
    n = len(rods["length"])
    eI = rng.choice(n, me, replace=False).astype(np.int32)
    eJ = ((eI + 1 + rng.integers(0, n - 2, me)) % n).astype(np.int32)
    c, t, L = rods["center"], rods["dir"], rods["length"]
    sI, sJ = rng.uniform(-0.5, 0.5, me) * L[eI], rng.uniform(-0.5, 0.5, me) * L[eJ]
    P, Q = c[eI] + sI[:, None] * t[eI], c[eJ] + sJ[:, None] * t[eJ]
    d = np.linalg.norm(P - Q, axis=1)
    kap = np.array([kappas[e % len(kappas)] for e in range(me)])
    bil = np.ones(me, np.int32)
    bil[-1] = 0
    d0 = d - free_len
    return dict(I=eI, J=eJ, gid_i=rods["gid"][eI], gid_j=rods["gid"][eJ], delta0=d0, nrm=(P - Q) / d[:, None],
                posI=sI[:, None] * t[eI], posJ=sJ[:, None] * t[eJ], labI=P, labJ=Q,
                gamma0=np.where(bil == 1, -kap * d0, 0.0), bilateral=bil, kappa=kap,
                gamma_lb=np.where(bil == 1, 0.0, -0.25), stress=np.arange(9.0 * me).reshape(me, 9) * 1e-3)


def test_caller_rows_bitexact():
    rods, cfg = systems.small_3d(n=500, box=3.6, seed=12)
    ext = make_ext(rods, 16, seed=3)
    ctx = capi.Context(**cfg)
    ctx.set_rods(rods)
    ctx.add_constraints(ext)
    rep = ctx.step()
    o = O.step(rods, O.cfg_from_dict(cfg), ext=ext)
    mc = o["n_collision"]
    g = ctx.constraint_arrays()
    assert rep["n_con"] == o["I"].size == mc + 16
    assert np.array_equal(g["I"], o["I"]) and np.array_equal(g["J"], o["J"])
    rw = o["rows"]
    for key in ("delta0", "nrm", "posI", "posJ", "rI", "rJ", "b"):
        assert_bits(g[key], rw[key], key)
    orp = o["report"]
    assert rep["iters"] == orp["iters"] and rep["matvecs"] == orp["matvecs"], (rep, orp)
    assert rep["fallback_used"] == orp["fallback_used"]
    assert_bits([rep["res0"], rep["res"]], [orp["res0"], orp["res"]], "residuals")
    assert_bits(ctx.gamma(), o["gamma"], "gamma")
    v = ctx.velocity()
    assert_bits(v["fcon"], o["fcon"], "F_con")
    assert_bits(v["u"], o["u"], "U")
    # caller blocks come back in order with this step's gamma; bilateral rows went negative (springs
    # stretched beyond the free length pull)
    blk = ctx.constraint_blocks()
    gam = ctx.gamma()
    for e in range(16):
        b = blk[mc + e]
        assert b.gid_i == ext["gid_i"][e] and b.gid_j == ext["gid_j"][e] and b.bilateral == ext["bilateral"][e]
        assert b.gamma == gam[mc + e] and b.gidx_i == ext["I"][e]
    assert np.any(gam[mc:mc + 15] < 0)
    # stress sums: collision rows from the virial of the unit force pair, caller rows from their matrices
    sc, se = ctx.con_stress()
    lab_i = np.array([list(b.lab_i) for b in blk[:mc]])
    lab_j = np.array([list(b.lab_j) for b in blk[:mc]])
    c = rods["center"]
    L = np.asarray(cfg["box_hi"], float) - np.asarray(cfg["box_lo"], float)
    assert_bits(lab_i, c[o["I"][:mc]] + rw["posI"][:3 * mc].reshape(-1, 3), "lab_i")
    assert_bits(lab_j, (c[o["J"][:mc]] + o["img"].reshape(-1, 3) * L) + rw["posJ"][:3 * mc].reshape(-1, 3), "lab_j")
    assert_bits(sc, O.stress_sum(rw["nrm"][:3 * mc], lab_i, lab_j, o["gamma"][:mc]), "collision stress")
    assert_bits(se, O.stress_sum(ext["nrm"], ext["labI"], ext["labJ"], o["gamma"][mc:], stress=ext["stress"],
                                 use_given=np.ones(16, np.uint8)), "caller stress")
    # the pool is consumed by the step: the next step has collision rows only
    ctx.euler()
    rep2 = ctx.step()
    c2, t2 = ctx.get_rods()
    assert rep2["n_con"] == ctx.num_constraints()
    o2 = O.step(dict(rods, center=c2, dir=t2), _cfg_step(cfg, 1))
    assert rep2["n_con"] == o2["I"].size
    assert_bits(ctx.gamma(), o2["gamma"], "gamma after the pool was consumed")


def _cfg_step(cfg, step):
    oc = O.cfg_from_dict(cfg)
    oc.step = step
    return oc


def test_caller_rows_fp32_form_and_warm_start_converge():
    """Caller rows with the FP32 form and with previous-gamma warm starts: FP64 residual <= tol."""
    rods, cfg = systems.small_3d(n=500, box=3.6, seed=12)
    for extra in (dict(precision=1), dict(warm_start=1)):
        ctx = capi.Context(**dict(cfg, **extra))
        ctx.set_rods(rods)
        for s in range(3):
            ctx.add_constraints(make_ext(rods, 10, seed=40 + s))
            rep = ctx.step()
            assert rep["converged"] == 1 and rep["res"] <= cfg["tol"], (extra, s, rep)
            ctx.euler()


def test_caller_rows_rejected_cases():
    rods, cfg = systems.small_3d(n=300, box=3.2, seed=2)
    ctx = capi.Context(**cfg)
    ctx.set_rods(rods)
    ext = make_ext(rods, 4, seed=1)
    bad = dict(ext, gid_j=np.array([10 ** 6] * 4))
    ctx.add_constraints(bad)
    with pytest.raises(capi.AlensError):
        ctx.step()
    ctx.clear_constraints()
    rep = ctx.step()  # the cleared pool leaves a plain collision step
    assert rep["converged"] == 1
