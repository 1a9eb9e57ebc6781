"""C ABI checks that need no GPU: every entry point declared in include/alens_capi.h is exported by
libalens.so, the ctypes mirrors have the C struct layouts, and the C++ adapter header compiles."""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

from alens_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "alens_capi.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(alens_[a-z_0-9]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 20
    L = ctypes.CDLL(capi.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert sorted(capi.EXPORTS) == names


def test_version_string():
    assert b"gfx950" in capi.lib().alens_version()


def test_struct_layouts_match_c():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "alens_capi.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(alens_config), offsetof(alens_config, seed),
         offsetof(alens_config, warm_start), sizeof(alens_con_block), offsetof(alens_con_block, stress),
         sizeof(alens_solve_report), offsetof(alens_solve_report, t_step_ms), sizeof(alens_tune_report),
         offsetof(alens_tune_report, best_ms), sizeof(alens_graph_info), offsetof(alens_graph_info, last_failure));
  return 0;
}'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "l.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "l")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        vals = [int(v) for v in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    want = [ctypes.sizeof(capi.Config), capi.Config.seed.offset, capi.Config.warm_start.offset,
            ctypes.sizeof(capi.ConBlock), capi.ConBlock.stress.offset, ctypes.sizeof(capi.Report),
            capi.Report.t_step_ms.offset, ctypes.sizeof(capi.TuneReport), capi.TuneReport.best_ms.offset,
            ctypes.sizeof(capi.GraphInfo), capi.GraphInfo.last_failure.offset]
    assert vals == want


def build_adapter_driver(outdir, name="adapter_step"):
    exe = os.path.join(outdir, name)
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "cpp", name + ".cpp"), "-o", exe,
                    "-L", os.path.dirname(capi.LIB_PATH), "-lalens", "-Wl,-rpath," + os.path.dirname(capi.LIB_PATH)],
                   check=True)
    return exe


def test_cpp_adapter_compiles_and_links():
    with tempfile.TemporaryDirectory() as d:
        exe = build_adapter_driver(d)
        assert os.path.exists(exe)


def test_cpp_rod_file_reader():
    """The C++ adapter's rod-file reader on the reference's inputs (group column and 'S' rods)."""
    golden = os.path.join(ROOT, "tests", "golden")
    with tempfile.TemporaryDirectory() as d:
        exe = build_adapter_driver(d, "adapter_restart")
        out = {}
        for f in ("dense_monolayer_TubuleInitial.dat", "mix_motor_sliding_TubuleInitial.dat"):
            p = subprocess.run([exe, "read", os.path.join(golden, f)], capture_output=True, text=True, timeout=60)
            assert p.returncode == 0, p.stdout + p.stderr
            tok = p.stdout.split()
            out[f] = dict(zip(tok[0::2], (int(v) for v in tok[1::2])))
    assert out["dense_monolayer_TubuleInitial.dat"] == dict(rods=9700, stationary=0, group0=-1)
    assert out["mix_motor_sliding_TubuleInitial.dat"] == dict(rods=2, stationary=1, group0=-1)


def test_copy_probe_rejects_bad_arguments():
    """Argument checks of alens_copy_probe come before any device work."""
    g = ctypes.c_double(-1.0)
    L = capi.lib()
    assert L.alens_copy_probe(0, 8, 1, ctypes.byref(g)) == capi.ALENS_E_ARG
    assert L.alens_copy_probe(0, 1 << 20, 0, ctypes.byref(g)) == capi.ALENS_E_ARG
    assert L.alens_copy_probe(0, 1 << 20, 1, None) == capi.ALENS_E_ARG
