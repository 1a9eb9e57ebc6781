"""bench.py's contract parts that need no GPU: the one-GPU refusal, the algorithmic byte count the
roofline object uses, and the committed traffic record it reads."""
import importlib.util
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert p.returncode == 2, (p.returncode, p.stdout, p.stderr)
    assert "one GPU" in (p.stdout + p.stderr)
    assert not any(line.startswith("{") for line in p.stdout.splitlines())


def test_algorithmic_bytes_per_matvec():
    """Per constraint: D and D^T each read 9 geometry words and the two rod ids, and the update
    reads x, g, g_prev, b and writes x/g; per rod: mobility read, U written and read back.  FP64:
    200 M + 144 N bytes; FP32: 108 M + 72 N."""
    b = _bench()
    assert b.bytes_per_matvec(1, 0) == 200 and b.bytes_per_matvec(0, 1) == 144
    assert b.bytes_per_matvec(1, 0, w=4) == 108 and b.bytes_per_matvec(0, 1, w=4) == 72
    assert b.bytes_per_matvec(328_000, 200_000) == 200 * 328_000 + 144 * 200_000


def test_committed_traffic_record_is_consistent():
    """One record per workload, keyed by the workload name the bench line carries; the default
    workload (the 1M-rod box) has one."""
    with open(os.path.join(ROOT, "profiles", "k_bbpgd_traffic.json")) as fh:
        allw = json.load(fh)
    assert "lognormal_3d" in allw
    for key, tr in allw.items():
        assert tr["workload"].split(":")[0] == key
        ratio = tr["traffic_bytes_per_matvec"] / tr["algorithmic_bytes_per_matvec"]
        assert abs(ratio - tr["traffic_over_algorithmic"]) < 1e-9
        assert 1.0 <= ratio < 3.0
        assert abs(tr["fetch_factor"] - 2.0) < 0.05 and abs(tr["write_factor"] - 1.0) < 0.05


def test_traffic_scales_with_this_runs_algorithmic_bytes():
    """roofline.traffic = measured/algorithmic ratio of the workload x this run's algorithmic bytes."""
    b = _bench()
    tr = b.load_traffic("lognormal_3d")
    assert tr is not None and b.load_traffic("no_such_workload") is None


def test_same_bits_compares_every_bit():
    b = _bench()
    import numpy as np

    x = np.array([1.0, -0.0, 2.5])
    assert b.same_bits(x, x.copy())
    assert not b.same_bits(x, np.array([1.0, 0.0, 2.5]))  # -0.0 and 0.0 differ in bits
    assert not b.same_bits(x, np.nextafter(x, 3.0))
    assert not b.same_bits(x, x[:2])
