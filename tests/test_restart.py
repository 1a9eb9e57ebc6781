"""Rod files, snapshots and restart (host side): the reference's rod text format read and written
back, the snapshot + restart record, and a restarted oracle trajectory that continues the
uninterrupted one."""
import os

import numpy as np
import pytest

from alens_amd import systems
from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _ends(r):
    c, d, L = r["center"], r["dir"], r["length"]
    return c - 0.5 * L[:, None] * d, c + 0.5 * L[:, None] * d


def test_reference_rod_file_round_trip(tmp_path):
    """The reference's DenseMonoLayer input (group column -1) written and read back: same rods."""
    a = systems.read_rod_file(os.path.join(GOLDEN, "dense_monolayer_TubuleInitial.dat"))
    assert a["gid"].size == 9700 and np.all(a["group"] == -1)
    p = tmp_path / "rods.dat"
    systems.write_rod_file(str(p), a, header_time=9.99999)
    b = systems.read_rod_file(str(p))
    for key in ("gid", "radius", "immovable", "group"):
        assert np.array_equal(a[key], b[key]), key
    for ea, eb in zip(_ends(a), _ends(b)):
        assert np.max(np.abs(ea - eb)) <= 4e-15 * np.max(np.abs(ea))
    lines = p.read_text().splitlines()
    assert lines[0] == "9700" and float(lines[1]) == 9.99999
    assert len(lines[2].split()) == 10 and lines[2].split()[-1] == "-1"


def test_stationary_rods_and_groups_survive(tmp_path):
    a = systems.read_rod_file(os.path.join(GOLDEN, "mix_motor_sliding_TubuleInitial.dat"))
    assert list(a["immovable"]) == [1, 0] and list(a["group"]) == [-1, -1]  # no group column there
    a["group"] = np.array([3, 7], np.int32)
    systems.write_snapshot(str(tmp_path), 0, a, step=0)
    b, info = systems.read_restart(str(tmp_path))
    assert list(b["immovable"]) == [1, 0] and list(b["group"]) == [3, 7]
    assert info["step"] == 1 and info["snap_id"] == 1
    assert open(tmp_path / "SylinderAscii_0.dat").read().splitlines()[2].startswith("S 0 ")


def test_restart_record_names_the_snapshot(tmp_path):
    rods, _ = systems.small_3d(n=40, seed=5)
    systems.write_snapshot(str(tmp_path), 2, rods, step=10, time=1e-3)
    systems.write_snapshot(str(tmp_path), 7, rods, step=35, time=3.5e-3)  # the record names the last one
    b, info = systems.read_restart(str(tmp_path))
    assert info == dict(step=36, snap_id=8, time=3.5e-3)
    assert np.array_equal(b["gid"], rods["gid"])
    assert np.max(np.abs(b["center"] - rods["center"])) <= 1e-15 * 4.0
    (tmp_path / systems.RESTART_FILE).write_text("snapID 1\nsnapFile SylinderAscii_2.dat\n")
    with pytest.raises(ValueError, match="stepCount"):
        systems.read_restart(str(tmp_path))


def test_restarted_oracle_trajectory_continues():
    """Steps 0..2 uninterrupted against steps 0..1, a snapshot, and step 2 from the restart: the same
    constraint pairs (by gid), the same iteration count, and gamma to 1e-8 relative (the restart
    recomputes centre and direction from the stored endpoints, a last-bit change)."""
    rods, cfg = systems.small_3d(n=1000, seed=3)
    oc = O.cfg_from_dict(cfg)
    cur = rods
    for s in range(2):
        oc.step = s
        o = O.step(cur, oc)
        c, t = O.euler(cur, o["u"], oc)
        cur = dict(cur, center=c, dir=t)
    import tempfile

    with tempfile.TemporaryDirectory() as d:
        systems.write_snapshot(d, 1, cur, step=1)
        back, info = systems.read_restart(d)
    assert info["step"] == 2
    oc.step = 2
    a = O.step(cur, oc)
    oc.step = info["step"]
    b = O.step(dict(back, immovable=None), oc)
    ga, gb = cur["gid"], back["gid"]
    pa = sorted(zip(ga[a["I"]].tolist(), ga[a["J"]].tolist(), a["gamma"].tolist()))
    pb = sorted(zip(gb[b["I"]].tolist(), gb[b["J"]].tolist(), b["gamma"].tolist()))
    assert [p[:2] for p in pa] == [p[:2] for p in pb] and len(pa) > 300
    assert a["report"]["iters"] == b["report"]["iters"]
    xa, xb = np.array([p[2] for p in pa]), np.array([p[2] for p in pb])
    assert np.max(np.abs(xa - xb)) <= 1e-8 * np.max(np.abs(xa))
