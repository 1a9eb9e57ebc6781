"""GPU parity: libalens.so (through the C ABI) against the FP64 CPU oracle on the same inputs.

Both sides evaluate the same operation sequence in IEEE binary64 with no contraction and the same
fixed reduction order, so the bar here is bit-for-bit: the pair set and every row field, the BBPGD
iteration count, residuals, gamma, U_nc, F_con and U, and the rod state after Euler steps.
"""
import json
import os

import numpy as np
import pytest

from alens_amd import capi, systems
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def assert_bits(a, b, what):
    a = np.asarray(a, np.float64).reshape(-1)
    b = np.asarray(b, np.float64).reshape(-1)
    assert a.shape == b.shape, f"{what}: shape {a.shape} vs {b.shape}"
    bad = np.nonzero(bits(a) != bits(b))[0]
    if bad.size:
        k = bad[0]
        raise AssertionError(f"{what}: {bad.size} of {a.size} differ; first at {k}: {a[k]!r} vs {b[k]!r}")


def gpu_step(rods, cfg, **extra):
    ctx = capi.Context(**dict(cfg, **extra))
    ctx.set_rods(rods)
    rep = ctx.step()
    out = ctx.constraint_arrays()
    out["gamma"] = ctx.gamma()
    out.update(ctx.velocity())
    out["report"] = rep
    return ctx, out


def compare_full(rods, cfg, g, o):
    assert np.array_equal(g["I"], o["I"]) and np.array_equal(g["J"], o["J"]), "pair set differs"
    assert np.array_equal(g["img"], o["img"]), "periodic images differ"
    rw = o["rows"]
    assert_bits(g["delta0"], rw["delta0"], "delta0")
    assert_bits(g["nrm"], rw["nrm"], "normal")
    assert_bits(g["posI"], rw["posI"], "posI")
    assert_bits(g["posJ"], rw["posJ"], "posJ")
    assert_bits(g["rI"], rw["rI"], "posI x n")
    assert_bits(g["rJ"], rw["rJ"], "posJ x n")
    assert_bits(g["mob"], o["mob"], "mobility")
    assert_bits(g["unc"], o["unc"], "U_nc")
    assert_bits(g["b"], rw["b"], "b")
    gr, orp = g["report"], o["report"]
    assert gr["iters"] == orp["iters"] and gr["matvecs"] == orp["matvecs"], (gr, orp)
    assert gr["converged"] == orp["converged"]
    assert_bits([gr["res0"], gr["res"]], [orp["res0"], orp["res"]], "residuals")
    assert_bits(g["gamma"], o["gamma"], "gamma")
    assert_bits(g["fcon"], o["fcon"], "F_con")
    assert_bits(g["u"], o["u"], "U")


def test_small_3d_step_bitexact():
    rods, cfg = systems.small_3d()
    ctx, g = gpu_step(rods, cfg)
    o = O.step(rods, O.cfg_from_dict(cfg))
    assert o["I"].size > 300
    compare_full(rods, cfg, g, o)
    assert g["report"]["converged"] == 1 and g["report"]["res"] <= cfg["tol"]


def test_small_3d_trajectory_bitexact():
    """Five full steps (constraints + Euler) stay bit-identical to the oracle's trajectory."""
    rods, cfg = systems.small_3d(n=600, box=3.4, seed=77)
    ctx = capi.Context(**cfg)
    ctx.set_rods(rods)
    oc = O.cfg_from_dict(cfg)
    cur = dict(rods)
    for step in range(5):
        oc.step = step
        o = O.step(cur, oc)
        ctx.step()
        g = ctx.constraint_arrays()
        g["gamma"] = ctx.gamma()
        g.update(ctx.velocity())
        assert np.array_equal(g["I"], o["I"]) and np.array_equal(g["J"], o["J"])
        assert_bits(g["gamma"], o["gamma"], f"gamma step {step}")
        assert_bits(g["u"], o["u"], f"U step {step}")
        ctx.euler()
        c, t = O.euler(cur, o["u"], oc)
        cur = dict(cur, center=c, dir=t)
        gc, gt = ctx.get_rods()
        assert_bits(gc, c, f"centers after step {step}")
        assert_bits(gt, t, f"directions after step {step}")
    assert ctx.step_count == 5


def test_dense_monolayer_example_bitexact():
    """The reference's DenseMonoLayer input (9,700 rods, 11 x 11 periodic, monolayer dynamics)."""
    rods = systems.read_rod_file(os.path.join(GOLDEN, "dense_monolayer_TubuleInitial.dat"))
    cfg = dict(box_lo=(0, 0, 0), box_hi=(11, 11, 11), pbc=(1, 1, 0), monolayer=1, col_buf=0.025, viscosity=1.0,
               kT=0.00411, dt=1e-5, tol=1e-6, max_ite=100000, seed=1234)
    ctx, g = gpu_step(rods, cfg)
    o = O.step(rods, O.cfg_from_dict(cfg))
    assert o["I"].size > 1000
    compare_full(rods, cfg, g, o)
    # monolayer dynamics: no out-of-plane motion
    u = g["u"].reshape(-1, 6)
    assert np.all(u[:, 2] == 0) and np.all(u[:, 3] == 0) and np.all(u[:, 4] == 0)


def test_dense_straining_example_bitexact():
    rods = systems.read_rod_file(os.path.join(GOLDEN, "dense_straining_TubuleInitial.dat"))
    cfg = dict(box_lo=(0, 0, 0), box_hi=(3, 1, 1), pbc=(1, 0, 0), monolayer=0, col_buf=0.025, viscosity=1.0,
               kT=-0.00411, dt=1e-4, tol=1e-5, max_ite=1000000, seed=1234)
    ctx, g = gpu_step(rods, cfg)
    o = O.step(rods, O.cfg_from_dict(cfg))
    compare_full(rods, cfg, g, o)


def test_flexible_filament_example_bitexact():
    """The reference's FlexibleFilament input: 200 beads (length 1e-6 < diameter 0.02: spheres) in a
    chain 0.0202 um apart, open 10 um box (Examples/FlexibleFilament/RunConfig.yaml).  Every
    neighbouring pair is a contact (gap 2e-4 um < colBuf 0.02)."""
    rods = systems.read_rod_file(os.path.join(GOLDEN, "flexible_filament_TubuleInitial.dat"))
    cfg = dict(box_lo=(-5, -5, -5), box_hi=(5, 5, 5), pbc=(0, 0, 0), monolayer=0, col_buf=0.02, viscosity=1.0,
               kT=0.00411, dt=1e-4, tol=1e-5, max_ite=1000000, seed=1234)
    ctx, g = gpu_step(rods, cfg)
    o = O.step(rods, O.cfg_from_dict(cfg))
    assert o["I"].size == 199 and np.all(o["J"] - o["I"] == 1)
    compare_full(rods, cfg, g, o)


def test_active_3d_nematics_like_bitexact():
    """Examples/Active3DNematics/RunConfig.yaml's system drawn here (its initial file is generated at
    run time): 500 rods, L = 0.25 um, D = 0.025 um, periodic 0.7 um box, no Brownian motion -- dense
    and overlapping at the start."""
    rng = np.random.default_rng(500)
    n = 500
    z = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    d = np.stack([np.sqrt(1 - z * z) * np.cos(ph), np.sqrt(1 - z * z) * np.sin(ph), z], 1)
    rods = dict(gid=np.arange(n, dtype=np.int32), center=rng.uniform(-0.35, 0.35, (n, 3)), dir=d,
                length=np.full(n, 0.25), radius=np.full(n, 0.0125), immovable=None)
    cfg = dict(box_lo=(-0.35, -0.35, -0.35), box_hi=(0.35, 0.35, 0.35), pbc=(1, 1, 1), monolayer=0, col_buf=0.025,
               viscosity=0.01, kT=-0.00411, dt=1e-4, tol=1e-5, max_ite=100000, seed=1234)
    ctx, g = gpu_step(rods, cfg)
    o = O.step(rods, O.cfg_from_dict(cfg))
    assert o["I"].size > 1000
    compare_full(rods, cfg, g, o)


def test_mix_motor_sliding_has_no_contacts():
    """Two parallel rods 0.07 um apart (gap 0.045 > colBuf 0.025): no constraint."""
    rods = systems.read_rod_file(os.path.join(GOLDEN, "mix_motor_sliding_TubuleInitial.dat"))
    cfg = dict(box_lo=(0, 0, 0), box_hi=(20, 1, 1), pbc=(0, 0, 0), col_buf=0.025, viscosity=1.0, kT=-0.00411,
               dt=1e-5, tol=1e-5, max_ite=10000)
    ctx, g = gpu_step(rods, cfg)
    assert g["I"].size == 0
    assert np.all(g["u"] == 0.0)  # no noise, no force, no contact
    assert g["report"]["converged"] == 1


@pytest.mark.parametrize("kind,n", [("monolayer_slab", 20000), ("lognormal_3d", 20000)])
def test_scaled_workloads_bitexact(kind, n):
    rods, cfg = systems.scaled(kind, n)
    ctx, g = gpu_step(rods, cfg)
    o = O.step(rods, O.cfg_from_dict(cfg))
    compare_full(rods, cfg, g, o)


def test_rerun_is_bit_identical():
    rods, cfg = systems.scaled("lognormal_3d", 30000, seed=5)
    _, a = gpu_step(rods, cfg)
    _, b = gpu_step(rods, cfg)
    assert np.array_equal(a["I"], b["I"])
    assert_bits(a["gamma"], b["gamma"], "gamma rerun")
    assert_bits(a["u"], b["u"], "U rerun")


def test_bcqp_solve_matches_exact_lcp():
    """Stand-alone BCQP entry on the 50-digit active-set solutions (tests/golden/lcp_small.json)."""
    cases = json.load(open(os.path.join(GOLDEN, "lcp_small.json")))["cases"]
    ctx = capi.Context(box_lo=(0, 0, 0), box_hi=(1, 1, 1), pbc=(0, 0, 0), tol=1e-13, max_ite=200000, dt=1e-4)
    for cs in cases:
        I = np.array(cs["I"], np.int32)
        J = np.array(cs["J"], np.int32)
        rw = {k: np.array(cs[k]) for k in ("nrm", "rI", "rJ", "b")}
        mob = np.array(cs["mob"])
        gam, rep = ctx.bcqp_solve(cs["n"], I, J, rw["nrm"], rw["rI"], rw["rJ"], rw["b"], mob)
        exact = np.array([float(x) for x in cs["gamma"]])
        assert np.max(np.abs(gam - exact)) <= 1e-9 * max(1.0, np.max(np.abs(exact))), cs
        og, orep = O.bbpgd(cs["n"], I, J, rw, mob, O.make_cfg(dt=cs["dt"], tol=1e-13, max_ite=200000))
        assert_bits(gam, og, "bcqp gamma vs oracle")
        assert rep["iters"] == orep["iters"]


def test_warm_start_previous_gamma_bitexact_and_fallback():
    """Warm start from the previous step's gamma: same iterate path as the oracle given the same x0."""
    rods, cfg = systems.small_3d(n=800, box=3.7, seed=11)
    ctx = capi.Context(**dict(cfg, warm_start=1))
    ctx.set_rods(rods)
    oc = O.cfg_from_dict(cfg)
    r0 = ctx.step()
    g0 = ctx.constraint_arrays()
    gam0 = ctx.gamma()
    o0 = O.step(rods, oc)
    assert_bits(gam0, o0["gamma"], "step 0 gamma")
    ctx.euler()
    c, t = O.euler(rods, o0["u"], oc)
    rods1 = dict(rods, center=c, dir=t)
    oc.step = 1
    r1 = ctx.step()
    g1 = ctx.constraint_arrays()
    # the warm start the library used: previous gamma matched on (I, J)
    key0 = {(int(i), int(j)): v for i, j, v in zip(g0["I"], g0["J"], gam0)}
    x0 = np.array([key0.get((int(i), int(j)), 0.0) for i, j in zip(g1["I"], g1["J"])])
    mob = O.mobility(rods1, oc)
    un = O.unc(rods1, mob, oc)
    I, J, img = O.pairs(rods1, oc)
    rw = O.rows(rods1, oc, I, J, img, un)
    og, orep = O.bbpgd(len(rods1["length"]), I, J, rw, mob, oc, x0=x0)
    assert_bits(ctx.gamma(), og, "warm gamma")
    assert r1["iters"] == orep["iters"] and r1["fallback_used"] == orep["fallback_used"]
    assert r1["iters"] < r0["iters"]


def test_error_paths():
    rods, cfg = systems.small_3d(n=100, box=3.2)
    ctx = capi.Context(**cfg)
    with pytest.raises(capi.AlensError) as e:
        ctx.step()
    assert e.value.status == capi.ALENS_E_STATE
    bad = dict(rods, center=rods["center"].copy())
    bad["center"][3, 1] = np.nan
    with pytest.raises(capi.AlensError) as e:
        ctx.set_rods(bad)
    assert e.value.status == capi.ALENS_E_NONFINITE
    tiny = capi.Context(**dict(cfg, box_hi=(2.0, 2.0, 2.0)))
    with pytest.raises(capi.AlensError) as e:
        tiny.set_rods(rods)
    assert e.value.status == capi.ALENS_E_BOX
    # a 2-iteration cap reports NOCONV but still writes a result
    dense, dcfg = systems.small_3d(n=1000)
    c2 = capi.Context(**dict(dcfg, max_ite=2))
    c2.set_rods(dense)
    rep = c2.step()
    assert rep["status_code"] == capi.ALENS_E_NOCONV and rep["iters"] == 2 and rep["converged"] == 0
    o = O.step(dense, O.cfg_from_dict(dict(dcfg, max_ite=2)))
    assert_bits(c2.gamma(), o["gamma"], "gamma at the iteration cap")


def test_constraint_blocks_follow_reference_layout():
    rods, cfg = systems.small_3d(n=500, box=3.2, seed=3)
    ctx, g = gpu_step(rods, cfg)
    blocks = ctx.constraint_blocks()
    assert len(blocks) == g["I"].size
    for k, cb in enumerate(blocks[:50]):
        i, j = g["I"][k], g["J"][k]
        assert cb.gid_i == rods["gid"][i] and cb.gid_j == rods["gid"][j]
        n = np.array(cb.norm_i)
        assert np.allclose(np.array(cb.norm_j), -n)
        # lab points are the closest points; their difference is (delta0 + radii) along the normal
        d = np.array(cb.lab_i) - np.array(cb.lab_j)
        assert np.allclose(d, (cb.delta0 + 2 * rods["radius"][i]) * n, atol=1e-12)
        assert np.allclose(np.array(cb.lab_i) - rods["center"][i], np.array(cb.pos_i), atol=1e-12)
        assert cb.bilateral == 0 and cb.gamma_lb == 0.0 and cb.gamma == g["gamma"][k]


def _kkt_residual(n, I, J, rw, mob, cfg_o, gamma):
    """max |min(gamma, A gamma + b)| with A applied by the oracle's FP64 operators."""
    _, U = O.apply_D_M(n, I, J, rw, mob, cfg_o, gamma)
    g = cfg_o.dt * O.apply_DT(I, J, rw, U) + rw["b"]
    return float(np.max(np.abs(np.minimum(gamma, g)))) if gamma.size else 0.0


@pytest.mark.parametrize("kind,n", [("small_3d", 1000), ("lognormal_3d", 20000), ("monolayer_slab", 20000)])
def test_precision_form_meets_fp64_residual(kind, n):
    """FP32 iterations + FP64 residual/refinement: the returned gamma satisfies the FP64 KKT residual
    tol, and its error against a tight FP64 solve is of the same order as the plain FP64 method's."""
    rods, cfg = systems.scaled(kind, n) if kind != "small_3d" else systems.small_3d()
    ctx, g = gpu_step(rods, cfg, precision=1)
    rep = g["report"]
    assert rep["converged"] == 1 and rep["res"] <= cfg["tol"] and rep["iters_fp32"] > 0
    oc = O.cfg_from_dict(cfg)
    o = O.step(rods, oc)
    assert np.array_equal(g["I"], o["I"]) and np.array_equal(g["J"], o["J"])
    nn = len(rods["length"])
    assert _kkt_residual(nn, o["I"], o["J"], o["rows"], o["mob"], oc, g["gamma"]) <= cfg["tol"]
    tight, _ = O.bbpgd(nn, o["I"], o["J"], o["rows"], o["mob"], O.cfg_from_dict(dict(cfg, tol=1e-11)))
    scale = max(np.max(np.abs(tight)), 1e-300)
    err_form = np.max(np.abs(g["gamma"] - tight)) / scale
    err_plain = np.max(np.abs(o["gamma"] - tight)) / scale
    print(f"{kind}: gamma rel err vs tight FP64 solve: precision form {err_form:.2e}, plain FP64 {err_plain:.2e}; "
          f"iters fp32 {rep['iters_fp32']} + fp64 {rep['iters']} vs plain {o['report']['iters']}")
    assert err_form <= max(20 * err_plain, 1e-6)


def test_sort_rods_spatial_order_bitexact():
    """After the Morton reorder the step is still bit-identical to the oracle run on the permuted rods."""
    rods, cfg = systems.scaled("lognormal_3d", 20000, seed=9)
    ctx = capi.Context(**cfg)
    ctx.set_rods(rods)
    perm = ctx.sort_rods()
    assert np.array_equal(np.sort(perm), np.arange(20000))
    pr = {k: (v[perm] if isinstance(v, np.ndarray) else v) for k, v in rods.items()}
    c, t = ctx.get_rods()
    assert_bits(c, pr["center"], "permuted centers")
    # neighbours are close in memory: mean index distance of contact pairs drops sharply
    rep = ctx.step()
    g = ctx.constraint_arrays()
    g["gamma"] = ctx.gamma()
    g.update(ctx.velocity())
    g["report"] = rep
    o = O.step(pr, O.cfg_from_dict(cfg))
    compare_full(pr, cfg, g, o)
    blocks = ctx.constraint_blocks()
    assert all(b.gid_i == pr["gid"][b.gidx_i] for b in blocks[:100])
    o_unsorted = O.pairs(rods, O.cfg_from_dict(cfg))
    assert np.mean(np.abs(g["J"] - g["I"])) < 0.25 * np.mean(np.abs(o_unsorted[1] - o_unsorted[0]))


def test_batched_systems_match_single_system_oracle():
    """Several independent systems (3-D periodic, slab, monolayer example) in one context and one
    solver launch: every system is bit-identical to its own single-system oracle run; per-system
    iteration counts differ and are reported separately."""
    specs = [systems.small_3d(n=700, box=3.5, seed=41), systems.scaled("monolayer_slab", 6000, seed=42),
             systems.small_3d(n=300, box=3.2, seed=43), systems.scaled("lognormal_3d", 16000, seed=44)]
    # globally unique gids so that the replicas draw independent noise
    off = 0
    parts = []
    for rods, cfg in specs:
        r = dict(rods, gid=rods["gid"] + off)
        off += len(rods["length"])
        parts.append((r, cfg))
    base = specs[0][1]
    ctx = capi.Context(**base)
    offs = ctx.set_systems(parts)
    rep = ctx.step()
    g = ctx.constraint_arrays()
    gam = ctx.gamma()
    vel = ctx.velocity()
    sreps = ctx.system_reports()
    assert len(sreps) == 4
    for s, (r, cfg) in enumerate(parts):
        # one context = one set of physical settings (seed, dt, ...); only the box is per system
        o = O.step(r, O.cfg_from_dict(dict(base, box_lo=cfg["box_lo"], box_hi=cfg["box_hi"], pbc=cfg["pbc"],
                                           monolayer=cfg["monolayer"])))
        lo, hi = offs[s], offs[s + 1]
        sel = (g["I"] >= lo) & (g["I"] < hi)
        assert np.array_equal(g["I"][sel] - lo, o["I"]) and np.array_equal(g["J"][sel] - lo, o["J"])
        assert_bits(g["delta0"][sel], o["rows"]["delta0"], f"system {s} delta0")
        assert_bits(gam[sel], o["gamma"], f"system {s} gamma")
        assert_bits(vel["u"][6 * lo:6 * hi], o["u"], f"system {s} U")
        assert sreps[s]["iters"] == o["report"]["iters"], (s, sreps[s], o["report"])
    assert rep["iters"] == max(r["iters"] for r in sreps)
    # the batch keeps its per-system boxes through the Euler step
    ctx.euler()
    c, _ = ctx.get_rods()
    for s, (r, cfg) in enumerate(parts):
        cc = c[offs[s]:offs[s + 1]]
        for k in range(3):
            if cfg["pbc"][k]:
                assert np.all(cc[:, k] >= cfg["box_lo"][k]) and np.all(cc[:, k] < cfg["box_hi"][k])


@pytest.mark.parametrize("case", ["nine_small", "one_large"])
def test_many_systems_match_single_system_oracle(case):
    """More systems than reduction groups per block: with nine small systems the fold takes one
    wave per system; with a system of more than 1,024 chunks among them it keeps the 256-lane
    groups.  Either way every system is bit-identical to its own single-system oracle run."""
    if case == "nine_small":
        specs = [systems.small_3d(n=300 + 60 * q, box=3.5 + 0.1 * q, seed=70 + q) for q in range(6)]
        specs += [systems.scaled("monolayer_slab", 2500 + 500 * q, seed=80 + q) for q in range(3)]
    else:
        specs = [systems.scaled("lognormal_3d", 180000, seed=90)]
        specs += [systems.small_3d(n=400 + 100 * q, box=3.2, seed=91 + q) for q in range(4)]
    off = 0
    parts = []
    for rods, cfg in specs:
        parts.append((dict(rods, gid=rods["gid"] + off), cfg))
        off += len(rods["length"])
    base = specs[0][1]
    ctx = capi.Context(**base)
    offs = ctx.set_systems(parts)
    ctx.step()
    g = ctx.constraint_arrays()
    gam = ctx.gamma()
    vel = ctx.velocity()
    sreps = ctx.system_reports()
    ctx.close()
    nch = [(int(r["n_con"]) + 511) // 512 for r in sreps]
    assert len(sreps) == len(specs) > 4
    assert (max(nch) > 1024) == (case == "one_large"), nch
    for s, (r, cfg) in enumerate(parts):
        o = O.step(r, O.cfg_from_dict(dict(base, box_lo=cfg["box_lo"], box_hi=cfg["box_hi"], pbc=cfg["pbc"],
                                           monolayer=cfg["monolayer"])))
        lo, hi = offs[s], offs[s + 1]
        sel = (g["I"] >= lo) & (g["I"] < hi)
        assert np.array_equal(g["I"][sel] - lo, o["I"]) and np.array_equal(g["J"][sel] - lo, o["J"])
        assert_bits(gam[sel], o["gamma"], f"{case}: system {s} gamma")
        assert_bits(vel["u"][6 * lo:6 * hi], o["u"], f"{case}: system {s} U")
        assert sreps[s]["iters"] == o["report"]["iters"], (case, s, sreps[s], o["report"])


def test_autotune_keeps_results_bit_identical():
    """The autotuner times the last step's solve under candidate schedules and keeps the fastest; the
    step's outputs stay as they were, and later steps are bit-identical to an untuned context."""
    rods, cfg = systems.scaled("monolayer_slab", 20000, seed=5)
    for extra in (dict(), dict(precision=1)):
        a = capi.Context(**dict(cfg, **extra))
        b = capi.Context(**dict(cfg, **extra))
        for c in (a, b):
            c.set_rods(rods)
            c.sort_rods()
        a.step()
        b.step()
        g_before = a.gamma()
        tr = a.autotune()
        assert tr["candidates"] >= 6 and 0 < tr["best_ms"] <= tr["default_ms"], tr
        assert tr["entries"] in (0, 2) and tr["bar_mode"] in (0, 1, 2, 3) and tr["xcd_map"] in (0, 1), tr
        assert tr["group_width"] in (0, 128, 256), tr
        assert tr["stream_hint"] in (0, 1, 2), tr
        assert_bits(a.gamma(), g_before, "gamma after tuning")
        assert_bits(a.velocity()["u"], b.velocity()["u"], "U after tuning")
        for s in range(3):
            a.euler()
            b.euler()
            ra, rb = a.step(), b.step()
            assert ra["iters"] == rb["iters"], (extra, s, ra, rb)
            assert_bits(a.gamma(), b.gamma(), f"gamma, step {s + 1} after tuning {extra}")


def _mixed_rods(n, box, seed, sphere_frac=0.3, immovable_frac=0.2):
    """Rods and spheres (L < 2r), some immovable, uniform in a box: the edge cases of the mobility
    (sphere drag, zero mobility) and of the segment distance (point segments)."""
    rng = np.random.default_rng(seed)
    L = rng.uniform(0.3, 1.0, n)
    sph = rng.random(n) < sphere_frac
    L[sph] = rng.uniform(0.0, 0.02, sph.sum())  # shorter than the diameter 0.025: spheres
    z = rng.uniform(-1, 1, n)
    ph = rng.uniform(0, 2 * np.pi, n)
    d = np.stack([np.sqrt(1 - z * z) * np.cos(ph), np.sqrt(1 - z * z) * np.sin(ph), z], 1)
    rods = dict(gid=np.arange(n, dtype=np.int32), center=rng.uniform(0, box, (n, 3)), dir=d, length=L,
                radius=np.full(n, 0.0125), immovable=(rng.random(n) < immovable_frac).astype(np.uint8))
    return rods


@pytest.mark.parametrize("pbc", [(1, 1, 1), (0, 0, 0)])
def test_spheres_and_immovable_rods_bitexact(pbc):
    rods = _mixed_rods(1500, 3.0, seed=17)
    cfg = dict(systems.COMMON, box_lo=(0, 0, 0), box_hi=(3.0, 3.0, 3.0), pbc=pbc, monolayer=0)
    ctx, g = gpu_step(rods, cfg)
    o = O.step(rods, O.cfg_from_dict(cfg))
    assert o["I"].size > 100
    compare_full(rods, cfg, g, o)
    imm = rods["immovable"].astype(bool)
    u = g["u"].reshape(-1, 6)
    assert np.all(u[imm] == 0.0), "immovable rods must not move"


@pytest.mark.parametrize("pbc", [(1, 1, 1), (0, 0, 0), (1, 0, 1)])
def test_mixed_radii_pair_set_bitexact(pbc):
    """Radii from 0.01 to 0.06 um (spheres the largest): a contact's closest-point midpoint can lie
    farther than the smaller rod's own radius + colBuf/2 from it, so the cell list must grow every
    rod's tube by the system's largest radius.  Rods also leave the box on the open axes."""
    rng = np.random.default_rng(31)
    n = 1200
    rods = _mixed_rods(n, 3.2, seed=29, immovable_frac=0.0)
    rods["radius"] = rng.uniform(0.01, 0.06, n)
    if not pbc[2]:
        rods["center"][:, 2] = rng.uniform(-0.3, 3.5, n)  # past both faces of the open z axis
    cfg = dict(systems.COMMON, box_lo=(0, 0, 0), box_hi=(3.2, 3.2, 3.2), pbc=pbc, monolayer=0)
    ctx, g = gpu_step(rods, cfg)
    o = O.step(rods, O.cfg_from_dict(cfg))
    assert o["I"].size > 150
    compare_full(rods, cfg, g, o)
    # the oracle's own cell search against brute force on the same input
    oc = O.cfg_from_dict(cfg)
    I, J, img = O.pairs(rods, oc, "brute")
    assert np.array_equal(I, o["I"]) and np.array_equal(J, o["J"]) and np.array_equal(img, o["img"])


def test_heavy_rod_split_is_bit_identical(monkeypatch):
    """In FP32 the solver lets light lanes form the products of heavy rods' later adjacency entries and
    the owners add them in order: a trajectory with that split equals one without it bit for bit."""
    rods, cfg = systems.scaled("monolayer_slab", 20000, seed=4)
    runs = []
    for sp in ("1", "0"):
        monkeypatch.setenv("ALENS_SOLVER_SPLIT", sp)
        ctx = capi.Context(**dict(cfg, precision=1))
        ctx.set_rods(rods)
        ctx.sort_rods()
        its, gam, deg = [], [], 0
        for _ in range(3):
            r = ctx.step()
            a = ctx.constraint_arrays()
            deg = max(deg, int((np.bincount(a["I"], minlength=20000) + np.bincount(a["J"], minlength=20000)).max()))
            its.append((r["iters_fp32"], r["iters"]))
            gam.append(ctx.gamma())
            u = ctx.velocity()["u"]
            ctx.euler()
        runs.append((its, gam, u, deg))
    assert runs[0][3] > 9, "the system must hold rods heavy enough to be split"
    assert runs[0][0] == runs[1][0]
    for ga, gb in zip(runs[0][1], runs[1][1]):
        assert_bits(ga, gb, "gamma with and without the split")
    assert_bits(runs[0][2], runs[1][2], "U with and without the split")


def test_external_force_and_monolayer_projection_bitexact():
    """A non-null F_ext (6 per rod) enters U_nc = M F_ext + U_B, and monolayer = 1 projects it."""
    rods, cfg = systems.small_3d(n=800, box=3.6, seed=23)
    rng = np.random.default_rng(4)
    f_ext = rng.normal(size=6 * 800) * 0.05
    for mono in (0, 1):
        c = dict(cfg, monolayer=mono)
        ctx = capi.Context(**c)
        ctx.set_rods(rods)
        rep = ctx.step(f_ext=f_ext)
        g = ctx.constraint_arrays()
        g["gamma"] = ctx.gamma()
        g.update(ctx.velocity())
        g["report"] = rep
        o = O.step(rods, O.cfg_from_dict(c), f_ext=f_ext)
        compare_full(rods, c, g, o)
        if mono:
            u = g["u"].reshape(-1, 6)
            assert np.all(u[:, 2] == 0.0) and np.all(u[:, 3] == 0.0) and np.all(u[:, 4] == 0.0)


def test_tiny_systems():
    """One rod (no constraints) and two touching rods (one constraint): the solver's trivial cases."""
    cfg = dict(systems.COMMON, box_lo=(0, 0, 0), box_hi=(4, 4, 4), pbc=(1, 1, 1), monolayer=0)
    one = dict(gid=np.array([7], np.int32), center=np.array([[1.0, 1.0, 1.0]]), dir=np.array([[0.0, 0.0, 1.0]]),
               length=np.array([1.0]), radius=np.array([0.0125]), immovable=None)
    ctx, g = gpu_step(one, cfg)
    assert g["report"]["n_con"] == 0 and g["report"]["converged"] == 1 and g["report"]["iters"] == 0
    o = O.step(one, O.cfg_from_dict(cfg))
    assert_bits(g["u"], o["u"], "one rod U")
    two = dict(gid=np.array([0, 1], np.int32), center=np.array([[1.0, 1.0, 1.0], [1.02, 1.0, 1.0]]),
               dir=np.array([[0.0, 0.0, 1.0], [0.0, 1.0, 0.0]]), length=np.array([1.0, 1.0]),
               radius=np.array([0.0125, 0.0125]), immovable=None)
    quiet = dict(cfg, kT=-1.0)  # no Brownian kick, which could open the 5 nm overlap by itself
    ctx, g = gpu_step(two, quiet)
    o = O.step(two, O.cfg_from_dict(quiet))
    assert g["report"]["n_con"] == 1
    compare_full(two, quiet, g, o)
    assert g["gamma"][0] > 0.0  # overlapping (gap -0.005 um): pushed apart


@pytest.mark.parametrize("max_blocks", [0, 3])
def test_barrier_variants_bit_identical(monkeypatch, max_blocks):
    """Every grid-barrier variant of the persistent solver (flat, grouped, grouped with longer polls,
    grouped by XCC with one L2 write-back per XCC) gives the same iterates, in plain steps and in a
    replayed graph; a grid of 3 blocks leaves most XCCs without blocks."""
    rods, cfg = systems.scaled("monolayer_slab", 50000, seed=11)
    runs = {}
    for mode in (1, 0, 2, 3):
        monkeypatch.setenv("ALENS_BARRIER_MODE", str(mode))
        monkeypatch.setenv("ALENS_SOLVER_MAX_BLOCKS", str(max_blocks))
        c = capi.Context(**cfg)
        c.set_rods(rods)
        c.sort_rods()
        out = []
        for s in range(4):
            if s == 1:
                c.set_graph(True)
            r = c.step()
            out.append((r["iters"], c.gamma().copy(), c.velocity()["u"].copy()))
            c.euler()
        if max_blocks == 0:
            assert c.graph_info()["replays"] >= 1, c.graph_info()
        runs[mode] = out
    for mode in (0, 2, 3):
        for s, (ra, rb) in enumerate(zip(runs[1], runs[mode])):
            assert ra[0] == rb[0] and ra[0] > 10, (mode, s, ra[0], rb[0])
            assert_bits(ra[1], rb[1], f"gamma, barrier mode {mode}, step {s}")
            assert_bits(ra[2], rb[2], f"U, barrier mode {mode}, step {s}")


def test_barrier_variants_bit_identical_replicas(monkeypatch):
    """The same across barrier variants for 12 replica systems solved in one launch (systems stop at
    different iterations while the others go on)."""
    parts, cfg = systems.replicas(nsys=12, nmin=2000, nmax=12000, seed=21)
    runs = {}
    for mode in (1, 0, 3):
        monkeypatch.setenv("ALENS_BARRIER_MODE", str(mode))
        c = capi.Context(**cfg)
        c.set_systems(parts)
        c.sort_rods()
        out = []
        for s in range(3):
            if s == 1:
                c.set_graph(True)
            c.step()
            out.append(([r["iters"] for r in c.system_reports()], c.gamma().copy(), c.velocity()["u"].copy()))
            c.euler()
        runs[mode] = out
    for mode in (0, 3):
        for s, (ra, rb) in enumerate(zip(runs[1], runs[mode])):
            assert ra[0] == rb[0] and len(set(ra[0])) > 1, (mode, s, ra[0], rb[0])
            assert_bits(ra[1], rb[1], f"gamma, barrier mode {mode}, step {s}")
            assert_bits(ra[2], rb[2], f"U, barrier mode {mode}, step {s}")


def _schedule_axis_bit_identical(monkeypatch, precision, var, values):
    """Runs three steps of one 30k-rod slab and of 12 replicas in one launch under each value of the
    schedule variable `var`; iteration counts, gamma and U must agree bit for bit with values[0]."""
    single = systems.scaled("monolayer_slab", 30000, seed=13)
    parts, rcfg = systems.replicas(nsys=12, nmin=2000, nmax=10000, seed=23)
    for kind in ("single", "replicas"):
        runs = {}
        for v in values:
            monkeypatch.setenv(var, str(v))
            if kind == "single":
                rods, cfg = single
                c = capi.Context(**dict(cfg, precision=precision))
                c.set_rods(rods)
            else:
                c = capi.Context(**dict(rcfg, precision=precision))
                c.set_systems(parts)
            c.sort_rods()
            out = []
            for s in range(3):
                r = c.step()
                out.append((r["iters"], r["iters_fp32"], c.gamma().copy(), c.velocity()["u"].copy()))
                c.euler()
            runs[v] = out
            c.close()
        for v in values[1:]:
            for s, (ra, rb) in enumerate(zip(runs[values[0]], runs[v])):
                assert ra[0] == rb[0] and ra[1] == rb[1], (kind, var, v, s, ra[:2], rb[:2])
                assert_bits(ra[2], rb[2], f"gamma, {kind}, {var}={v}, step {s}")
                assert_bits(ra[3], rb[3], f"U, {kind}, {var}={v}, step {s}")


@pytest.mark.parametrize("precision", [0, 1])
def test_group_width_variants_bit_identical(monkeypatch, precision):
    """The constraint phase's reduction groups of 256 or 128 lanes (a 128-lane group holds two virtual
    lanes per lane and folds them apart) give the fixed reduction order either way: iterates, gamma
    and U agree bit for bit under ALENS_GROUP_WIDTH = 0 (automatic), 256 and 128, for one system and
    for 12 replicas solved in one launch."""
    _schedule_axis_bit_identical(monkeypatch, precision, "ALENS_GROUP_WIDTH", (0, 256, 128))


@pytest.mark.parametrize("precision", [0, 1])
def test_stream_hint_bit_identical(monkeypatch, precision):
    """The cache hint of the constraint phase's row streams changes only where lines are kept: results
    agree bit for bit under ALENS_STREAM_HINT = 0 (automatic), 1 (cached) and 2 (non-temporal)."""
    _schedule_axis_bit_identical(monkeypatch, precision, "ALENS_STREAM_HINT", (0, 1, 2))


def test_copy_probe_rate():
    """alens_copy_probe reports a plausible device copy rate (read + write bytes per second)."""
    gbs = capi.copy_probe(256 << 20, reps=3)
    assert 500.0 < gbs < 20000.0, gbs
