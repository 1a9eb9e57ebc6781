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
