"""Host check of the colour-block field layout arithmetic shared by every
kernel (tests/layout_check.cpp against ollipticid_amd/csrc/kernels.hpp):
pidx is one to one, the stencil neighbour offsets, the colour-pass and
whole-row slot formulas and the transfers' k step all agree with pidx."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)


@pytest.mark.skipif(HIPCC is None, reason="hipcc not available")
def test_layout_arithmetic(tmp_path):
    exe = tmp_path / "layout_check"
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-I" + os.path.join(os.path.dirname(HERE), "include"),
                    "-o", str(exe), os.path.join(HERE, "layout_check.cpp")],
                   check=True, capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout


@pytest.mark.skipif(HIPCC is None, reason="hipcc not available")
def test_fused_sweep_indexing(tmp_path):
    """tests/fused_check.cpp: the fused double-buffered colour passes'
    tile rows, halo rows, colour slots and field choice per neighbour,
    replayed tile by tile in shuffled order, stay in bounds, read nothing
    another workgroup of the same launch writes, and give the bits of
    colour-by-colour sweeps."""
    exe = tmp_path / "fused_check"
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-I" + os.path.join(os.path.dirname(HERE), "include"),
                    "-o", str(exe), os.path.join(HERE, "fused_check.cpp")],
                   check=True, capture_output=True, timeout=300)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout
