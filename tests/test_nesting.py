"""Host check of the prolongation's nested fast path (tests/nesting_check.cpp
against ollipticid_amd/csrc/geometry.cpp): every linear transfer table that
nested_axis accepts interpolates, from its index map, exactly what its
nodes and weights do; the benchmark workload's tables are all nested."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CXX = shutil.which("g++")


@pytest.mark.skipif(CXX is None, reason="g++ not available")
def test_nested_prolongation_tables(tmp_path):
    exe = tmp_path / "nesting_check"
    subprocess.run([CXX, "-O1", "-std=c++17", "-I" + os.path.join(REPO, "include"), "-o", str(exe),
                    os.path.join(HERE, "nesting_check.cpp"),
                    os.path.join(REPO, "ollipticid_amd", "csrc", "geometry.cpp")],
                   check=True, capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout
    nested, tables = map(int, out.stdout.split()[1:3])
    assert nested >= 60 and tables == 72
