"""The C-ABI library loads and exports every symbol include/tdz.h declares
(no compute call: there is no GPU here), and bench.py keeps its contract."""
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "tdz.h")).read()
    return sorted(set(re.findall(r"\b(tdz_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    import tdsez_amd
    assert sorted(tdsez_amd.EXPORTS) == _declared()


def test_library_exports_every_symbol():
    import tdsez_amd
    if not os.path.exists(tdsez_amd.LIB_PATH):
        pytest.skip("libtdz.so not built")
    lib = tdsez_amd.lib()
    for name in _declared():
        assert hasattr(lib, name), name


def test_library_has_no_oracle_dependency():
    import tdsez_amd
    if not os.path.exists(tdsez_amd.LIB_PATH):
        pytest.skip("libtdz.so not built")
    out = subprocess.run(["nm", "-D", tdsez_amd.LIB_PATH], capture_output=True, text=True).stdout
    assert "orc_" not in out
    ldd = subprocess.run(["ldd", tdsez_amd.LIB_PATH], capture_output=True, text=True).stdout
    assert "liboracle" not in ldd


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
