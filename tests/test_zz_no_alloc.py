"""The allocation-free BCQP entry point (alens_bcqp_solve_static, its device memory one static array
of the library) against the allocating alens_bcqp_solve on the same rows: gamma, iteration and
matvec counts and residual bit-identical, also against the FP64 oracle's BBPGD, on two calls in a row
(the second smaller, so stale contents of the shared array would show), cold and warm started; sizes
beyond the array's limits are refused before any device work."""
import numpy as np
import pytest

from alens_amd import capi, systems
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _rows(kind, n, seed):
    rods, cfg = systems.scaled(kind, n, seed=seed)
    oc = O.cfg_from_dict(cfg)
    mob = O.mobility(rods, oc)
    un = O.unc(rods, mob, oc)
    I, J, img = O.pairs(rods, oc)
    rw = O.rows(rods, oc, I, J, img, un)
    return cfg, oc, I, J, rw, mob


def _bits(a):
    return np.ascontiguousarray(a, np.float64).view(np.uint64)


def test_static_solve_matches_allocating_solve():
    cases = (("lognormal_3d", 30000, 5), ("monolayer_slab", 8000, 6))
    ctx = None
    for kind, n, seed in cases:
        cfg, oc, I, J, rw, mob = _rows(kind, n, seed)
        m = I.size
        print(f"{kind}: {n} rods, {m} rows", flush=True)
        assert 1000 < m <= 131072
        if ctx is None:
            ctx = capi.Context(**cfg)
        args = (n, I, J, rw["nrm"], rw["rI"], rw["rJ"], rw["b"], mob)
        g_ref, r_ref = ctx.bcqp_solve(*args)
        g_st, r_st = ctx.bcqp_solve(*args, static=True)
        assert r_st["converged"] == 1 and r_st["status_code"] == capi.ALENS_OK
        assert np.array_equal(_bits(g_st), _bits(g_ref))
        for key in ("iters", "matvecs", "converged", "fallback_used", "n_con"):
            assert r_st[key] == r_ref[key], key
        assert _bits(np.array([r_st["res"]]))[0] == _bits(np.array([r_ref["res"]]))[0]
        g_o, r_o = O.bbpgd(n, I, J, rw, mob, oc)
        assert np.array_equal(_bits(g_st), _bits(g_o)) and r_o["iters"] == r_st["iters"]
        # warm start from a perturbed solution
        x0 = g_ref * 0.9 + 1e-7
        gw_ref, rw_ref = ctx.bcqp_solve(*args, x0=x0)
        gw_st, rw_st = ctx.bcqp_solve(*args, x0=x0, static=True)
        assert np.array_equal(_bits(gw_st), _bits(gw_ref)) and rw_st["iters"] == rw_ref["iters"]
        assert rw_st["fallback_used"] == rw_ref["fallback_used"]
    ctx.close()


def test_static_solve_refuses_sizes_beyond_its_array():
    ctx = capi.Context()
    mob = np.zeros(6 * 70000)
    mob[3::6] = 1.0
    mob[5::6] = 1.0

    def call(n, I, J):
        m = I.size
        nrm = np.tile([1.0, 0.0, 0.0], m)
        with pytest.raises(capi.AlensError) as e:
            ctx.bcqp_solve(n, I, J, nrm, np.zeros(3 * m), np.zeros(3 * m), np.ones(m), mob[:6 * n], static=True)
        assert e.value.status == capi.ALENS_E_ARG
        return str(e.value)

    big = 131073
    msg = call(2, np.zeros(big, np.int32), np.ones(big, np.int32))
    assert "131072 rows" in msg
    msg = call(65537, np.array([0], np.int32), np.array([1], np.int32))
    assert "65536 rods" in msg
    # 100000 rows all on rod 0: its slice alone would need 64 x 100000 padded slots
    k = 100000
    msg = call(65536, np.zeros(k, np.int32), (1 + np.arange(k) % 65535).astype(np.int32))
    assert "slots" in msg
    # the context still works after the refusals
    cfg, oc, I, J, rw, mob2 = _rows("small_3d", 500, 7)
    g, r = ctx.bcqp_solve(500, I, J, rw["nrm"], rw["rI"], rw["rJ"], rw["b"], mob2, static=True)
    assert r["converged"] == 1 and g.size == I.size
    ctx.close()
