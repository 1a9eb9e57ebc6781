"""The benchmark workloads at their full sizes (the 200,000-rod monolayer slab and the 1,000,000-rod
lognormal 3-D box), first step from the seeded start, checked through properties that need no
oracle solve at that size:
- the pair set, periodic images and every row field (delta0, normal, posI x n, posJ x n, b) equal
  the oracle's cell-list pairs and rows bit for bit (rows from the GPU's U_nc, which the smaller
  parity tests pin to the oracle's);
- the returned gamma satisfies the KKT conditions of the BCQP to the solver tolerance, with the
  gradient A gamma + b recomputed from the oracle's D, M and D^T;
- the end-of-step separations delta0 + dt D^T U from the GPU's own write-back U are non-negative to
  the tolerance, and zero where gamma > 0;
- the contact impulses balance: the forces of F_con = D gamma sum to zero over the rods.
"""
import numpy as np
import pytest

from alens_amd import capi, systems
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a, np.float64).reshape(-1).view(np.uint64)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("kind", ["monolayer_slab", "lognormal_3d"])
def test_full_size_pairs_rows_kkt_and_balance(kind, precision):
    """precision 1 is the FP32 form: FP32 iterations, then the FP64 residual check and FP64 refinement,
    so its gamma must meet the same FP64 KKT conditions."""
    rods, cfg = getattr(systems, kind)()
    n = rods["length"].size
    ctx = capi.Context(**dict(cfg, precision=precision))
    ctx.set_rods(rods)
    perm = ctx.sort_rods()
    rods = {k: (v[perm] if isinstance(v, np.ndarray) else v) for k, v in rods.items()}
    rep = ctx.step()
    assert rep["converged"] == 1 and rep["res"] <= cfg["tol"], rep
    g = ctx.constraint_arrays()
    vel = ctx.velocity()
    gam = ctx.gamma()
    m = g["I"].size
    print(f"{kind} precision {precision}: {n} rods, {m} constraints, {rep['iters']} FP64 and "
          f"{rep['iters_fp32']} FP32 iterations", flush=True)

    oc = O.cfg_from_dict(cfg)
    I, J, img = O.pairs(rods, oc)
    assert np.array_equal(g["I"], I) and np.array_equal(g["J"], J), "pair set differs from the oracle's"
    assert np.array_equal(g["img"], img), "periodic images differ"
    rw = O.rows(rods, oc, I, J, img, vel["unc"])
    for f in ("delta0", "nrm", "rI", "rJ", "b"):
        assert np.array_equal(_bits(g[f]), _bits(rw[f])), f"{f} differs from the oracle's rows"

    # KKT of min 1/2 x'Ax + b'x, x >= 0: x >= 0, A x + b >= -tol, |min(x, A x + b)| <= tol
    _, U = O.apply_D_M(n, I, J, rw, vel["mob"], oc, gam)
    grad = oc.dt * O.apply_DT(I, J, rw, U) + rw["b"]
    slack = cfg["tol"] * (1 + 1e-9)
    assert np.all(gam >= 0.0)
    assert grad.min() >= -slack, grad.min()
    assert np.max(np.abs(np.minimum(gam, grad))) <= slack
    assert np.count_nonzero(gam) > m // 10

    # the write-back: with the GPU's own U = U_nc + M D gamma, the separations at the end of the
    # step, delta0 + dt D^T U, are the gradient above: non-negative to the tolerance, and zero (to
    # the tolerance) where gamma > 0
    sep = rw["delta0"] + oc.dt * O.apply_DT(I, J, rw, vel["u"])
    assert np.max(np.abs(sep - grad)) <= 1e-9 * (1.0 + np.max(np.abs(grad)))
    assert sep.min() >= -slack
    assert np.max(np.abs(sep[gam > 0])) <= slack

    # Newton's third law: every contact puts opposite forces on its two rods
    f = vel["fcon"].reshape(-1, 6)[:, :3]
    assert np.all(np.abs(f.sum(0)) <= 1e-9 * np.abs(f).sum(0))


def test_full_size_replicas_per_system():
    """The 64-replica batch (1.13M rods, boxes and slabs of 2k-50k rods) in one context: for every
    system, its pairs and rows equal the oracle's for that system alone, its gamma satisfies the KKT
    conditions, its forces balance, and its own report says converged."""
    parts, cfg = systems.replicas()
    ctx = capi.Context(**cfg)
    offs = ctx.set_systems(parts)
    perm = ctx.sort_rods()  # within each system
    cat = {k: np.concatenate([np.asarray(r[k]).reshape(len(r["length"]), -1) for r, _ in parts])
           for k in ("gid", "center", "dir", "length", "radius")}
    cat = {k: (v[perm].reshape(-1) if v.shape[1] == 1 else v[perm]) for k, v in cat.items()}
    ctx.step()
    g = ctx.constraint_arrays()
    vel = ctx.velocity()
    gam = ctx.gamma()
    sreps = ctx.system_reports()
    assert len(sreps) == 64
    checked = 0
    for s, (_, box) in enumerate(parts):
        lo, hi = int(offs[s]), int(offs[s + 1])
        assert sreps[s]["converged"] == 1 and sreps[s]["res"] <= cfg["tol"], (s, sreps[s])
        rods = {k: v[lo:hi] for k, v in cat.items()}
        rods["immovable"] = None
        oc = O.cfg_from_dict(dict(cfg, **box))
        I, J, img = O.pairs(rods, oc)
        sel = (g["I"] >= lo) & (g["I"] < hi)
        assert np.array_equal(g["I"][sel] - lo, I) and np.array_equal(g["J"][sel] - lo, J), f"system {s} pairs"
        assert np.array_equal(g["img"][sel], img), f"system {s} images"
        rw = O.rows(rods, oc, I, J, img, vel["unc"][6 * lo:6 * hi])
        for f in ("delta0", "nrm", "rI", "rJ", "b"):
            assert np.array_equal(_bits(g[f][sel]), _bits(rw[f])), f"system {s}: {f}"
        n = hi - lo
        _, U = O.apply_D_M(n, I, J, rw, vel["mob"][6 * lo:6 * hi], oc, gam[sel])
        grad = oc.dt * O.apply_DT(I, J, rw, U) + rw["b"]
        slack = cfg["tol"] * (1 + 1e-9)
        assert np.all(gam[sel] >= 0.0) and grad.min() >= -slack, s
        assert np.max(np.abs(np.minimum(gam[sel], grad))) <= slack, s
        f = vel["fcon"].reshape(-1, 6)[lo:hi, :3]
        assert np.all(np.abs(f.sum(0)) <= 1e-9 * np.abs(f).sum(0) + 1e-300), s
        checked += I.size
    assert checked == g["I"].size


def _relaxed(kind, relax, **over):
    """A context on the full-size workload after `relax` graph-mode steps from the seeded start, and
    the rod arrays in the context's (spatial) order."""
    rods, cfg = getattr(systems, kind)()
    cfg = dict(cfg, **over)
    ctx = capi.Context(**cfg)
    ctx.set_rods(rods)
    perm = ctx.sort_rods()
    rods = {k: (v[perm] if isinstance(v, np.ndarray) else v) for k, v in rods.items()}
    ctx.set_graph(True)
    for _ in range(relax):
        ctx.step()
        ctx.euler()
    return ctx, rods, cfg


@pytest.mark.parametrize("kind,relax", [("monolayer_slab", 20), ("lognormal_3d", 8)])
def test_full_size_relaxed_step_bitexact(kind, relax):
    """At a relaxed state of the full-size workloads (after graph-mode steps from the seeded start),
    one more GPU step (a graph replay) and the FP64 oracle step from the same state agree bit for
    bit: pair set, iteration count, gamma, F_con and U."""
    ctx, rods, cfg = _relaxed(kind, relax)
    c, t = ctx.get_rods()
    st = ctx.step_count
    rep = ctx.step()
    assert ctx.graph_info()["replays"] > 0
    g = ctx.constraint_arrays()
    gam = ctx.gamma()
    vel = ctx.velocity()
    oc = O.cfg_from_dict(cfg)
    oc.step = st
    ref = O.step(dict(rods, center=c, dir=t), oc)
    print(f"{kind} step {st}: {g['I'].size} constraints, GPU {rep['iters']} / oracle {ref['report']['iters']} "
          f"iterations", flush=True)
    assert np.array_equal(g["I"], ref["I"]) and np.array_equal(g["J"], ref["J"])
    assert rep["iters"] == ref["report"]["iters"] and rep["converged"] == 1
    assert np.array_equal(_bits(gam), _bits(ref["gamma"])), "gamma differs"
    assert np.array_equal(_bits(vel["fcon"]), _bits(ref["fcon"])), "F_con differs"
    assert np.array_equal(_bits(vel["u"]), _bits(ref["u"])), "U differs"
    ctx.close()


def test_monolayer_500_steps_graph():
    """The 200,000-rod monolayer slab as its configuration states it: 500 FP64 steps, each one
    replay of the captured step graph.  Every step: converged to tol, the KKT conditions of the
    returned gamma (gradient from the oracle's D, M, D^T on the step's own rows) and zero net contact
    force.  On 10 sampled steps (every 50th), teacher-forced from the GPU's state before the step:
    pair set, iteration count, gamma and U bit-identical to the FP64 oracle's step."""
    ctx, rods, cfg = _relaxed("monolayer_slab", 0)
    oc = O.cfg_from_dict(cfg)
    n = rods["length"].size
    slack = cfg["tol"] * (1 + 1e-9)
    sampled = 0
    for s in range(500):
        check = s > 0 and s % 50 == 0 or s == 499
        if check:
            c, t = ctx.get_rods()
            st = ctx.step_count
        rep = ctx.step()
        assert rep["converged"] == 1 and rep["res"] <= cfg["tol"], (s, rep)
        g = ctx.constraint_arrays()
        gam = ctx.gamma()
        vel = ctx.velocity()
        rw = dict(nrm=g["nrm"].reshape(-1), rI=g["rI"].reshape(-1), rJ=g["rJ"].reshape(-1))
        _, U = O.apply_D_M(n, g["I"], g["J"], rw, vel["mob"], oc, gam)
        grad = oc.dt * O.apply_DT(g["I"], g["J"], rw, U) + g["b"]
        assert np.all(gam >= 0.0) and grad.min() >= -slack, s
        assert np.max(np.abs(np.minimum(gam, grad))) <= slack, s
        f = vel["fcon"].reshape(-1, 6)[:, :3]
        assert np.all(np.abs(f.sum(0)) <= 1e-9 * np.abs(f).sum(0) + 1e-300), s
        if check:
            ocs = O.cfg_from_dict(cfg)
            ocs.step = st
            ref = O.step(dict(rods, center=c, dir=t), ocs)
            assert np.array_equal(g["I"], ref["I"]) and np.array_equal(g["J"], ref["J"]), s
            assert rep["iters"] == ref["report"]["iters"], s
            assert np.array_equal(_bits(gam), _bits(ref["gamma"])), s
            assert np.array_equal(_bits(vel["u"]), _bits(ref["u"])), s
            sampled += 1
            print(f"step {s}: {g['I'].size} constraints, {rep['iters']} iterations, bit-identical to the oracle",
                  flush=True)
        ctx.euler()
    assert sampled == 10
    info = ctx.graph_info()
    assert info["replays"] >= 490 and info["capture_failures"] == 0, info
    ctx.close()
