"""The C++ call surface (SylinderSystem / ConstraintCollector pool / runStep / stepEuler) driving
libalens.so on the GPU."""
import subprocess
import tempfile

import pytest

from tests.test_capi_cpu import build_adapter_driver

pytestmark = pytest.mark.gpu


def test_cpp_adapter_runs_steps():
    with tempfile.TemporaryDirectory() as d:
        exe = build_adapter_driver(d)
        p = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        assert p.returncode == 0, p.stdout + p.stderr
        tok = p.stdout.split()
        vals = dict(zip(tok[0::2], tok[1::2]))
        assert int(vals["M"]) > 50 and vals["converged"] == "1" and vals["steps"] == "3", p.stdout
        # three stretched springs from the pool come back bilateral, pulling, with trace(stress sum)
        # = 3 * sum of their gammas (unit matrices given)
        assert vals["pooled"] == "3" and vals["pull"] == "1" and float(vals["trace"]) < 0, p.stdout


def test_cpp_adapter_restart(tmp_path):
    """writeResult after three steps, then reinitialize from TimeStepInfo.txt: the resumed system's
    next step has the same contact pairs and iteration count, and gamma to 1e-8 relative."""
    with tempfile.TemporaryDirectory() as d:
        exe = build_adapter_driver(d, "adapter_restart")
        p = subprocess.run([exe, "run", str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stdout + p.stderr
    tok = p.stdout.split()
    v = dict(zip(tok[0::2], tok[1::2]))
    assert int(v["M"]) > 50 and v["same_pairs"] == "1", p.stdout
    assert v["iters_a"] == v["iters_b"] and float(v["rel"]) <= 1e-8, p.stdout
    assert v["snap"] == "1" and v["step"] == "3", p.stdout
    assert (tmp_path / "SylinderAscii_0.dat").exists() and (tmp_path / "TimeStepInfo.txt").exists()
