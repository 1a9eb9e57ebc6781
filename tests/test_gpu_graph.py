"""Graph mode: after one plain step sizes the buffers, each step is one replay of a captured HIP graph
with every count kept on the device.  The results must be bit-identical to plain steps, for one system
and for a batch, cold and warm-started, FP64 and the FP32 form, and when a replay outgrows its
capacities (redone plainly)."""
import numpy as np
import pytest

from alens_amd import capi, systems

from .test_gpu_parity import assert_bits

pytestmark = pytest.mark.gpu


def _pair(setup, cfg, extra):
    a = capi.Context(**dict(cfg, **extra))
    b = capi.Context(**dict(cfg, **extra))
    for c in (a, b):
        setup(c)
        c.sort_rods()
    a.set_graph(True)
    return a, b


def _run_and_compare(a, b, steps, what):
    for s in range(steps):
        ra, rb = a.step(), b.step()
        assert ra["n_con"] == rb["n_con"] and ra["iters"] == rb["iters"], (what, s, ra, rb)
        assert ra["iters_fp32"] == rb["iters_fp32"] and ra["converged"] == rb["converged"], (what, s)
        # replays time collection and solve from clock readings inside the graph
        assert 0.0 < ra["t_collect_ms"] and 0.0 <= ra["t_solve_ms"] <= ra["t_step_ms"], (what, s, ra)
        assert ra["t_solve_ms"] > 0.0 or ra["n_con"] == 0, (what, s, ra)
        assert_bits([ra["res0"], ra["res"]], [rb["res0"], rb["res"]], f"{what} step {s} residuals")
        assert_bits(a.gamma(), b.gamma(), f"{what} step {s} gamma")
        va, vb = a.velocity(), b.velocity()
        assert_bits(va["u"], vb["u"], f"{what} step {s} U")
        ga, gb = a.constraint_arrays(), b.constraint_arrays()
        assert np.array_equal(ga["I"], gb["I"]) and np.array_equal(ga["J"], gb["J"])
        assert_bits(ga["nrm"], gb["nrm"], f"{what} step {s} normals")
        assert_bits(ga["b"], gb["b"], f"{what} step {s} b")
        sa, sb = a.system_reports(), b.system_reports()
        assert [r["iters"] for r in sa] == [r["iters"] for r in sb], (what, s)
        a.euler()
        b.euler()
    ca, cb = a.get_rods()[0], b.get_rods()[0]
    assert_bits(ca, cb, f"{what} positions")


@pytest.mark.parametrize("extra", [dict(), dict(warm_start=1), dict(precision=1, warm_start=1)])
def test_graph_steps_match_plain_steps(extra):
    rods, cfg = systems.scaled("monolayer_slab", 20000, seed=3)
    a, b = _pair(lambda c: c.set_rods(rods), cfg, extra)
    _run_and_compare(a, b, 6, f"slab {extra}")
    info = a.graph_info()
    assert info["captured"] == 1 and info["replays"] == 5, info


def test_graph_steps_batched_systems():
    parts, cfg = systems.replicas(nsys=6, nmin=2500, nmax=6000, seed=11)
    a, b = _pair(lambda c: c.set_systems(parts), cfg, dict(warm_start=1))
    _run_and_compare(a, b, 5, "replicas")
    assert a.graph_info()["replays"] == 4


def test_graph_replay_past_capacity_falls_back(monkeypatch):
    """With capacities below the counts every replay overflows; the guards keep the replay inside its
    buffers, and the step is redone plainly with identical results."""
    monkeypatch.setenv("ALENS_GRAPH_HEADROOM", "0.5")
    rods, cfg = systems.small_3d(n=1500, box=4.6, seed=8)
    a, b = _pair(lambda c: c.set_rods(rods), cfg, dict(warm_start=1))
    _run_and_compare(a, b, 4, "overflowing replays")
    assert a.graph_info()["replays"] == 0


def test_graph_steps_without_contacts():
    """Replays of steps with no contact at all (two parallel rods 0.07 um apart, no noise) and of a
    small system whose contacts come and go stay identical to plain steps."""
    import os

    golden = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    rods = systems.read_rod_file(os.path.join(golden, "mix_motor_sliding_TubuleInitial.dat"))
    cfg = dict(box_lo=(0, 0, 0), box_hi=(20, 1, 1), pbc=(0, 0, 0), col_buf=0.025, viscosity=1.0, kT=-0.00411,
               dt=1e-5, tol=1e-5, max_ite=10000)
    a, b = _pair(lambda c: c.set_rods(rods), cfg, dict())
    _run_and_compare(a, b, 4, "no contacts")
    assert a.graph_info()["replays"] == 3
    rods, cfg = systems.small_3d(n=60, box=4.0, seed=2)  # sparse: a handful of contacts, some steps none
    a, b = _pair(lambda c: c.set_rods(rods), cfg, dict(warm_start=1))
    _run_and_compare(a, b, 6, "sparse")
