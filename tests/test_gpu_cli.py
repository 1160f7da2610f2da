"""The drop-in command line on the GPU: same flags, same output files."""
import os
import subprocess

import numpy as np
import pytest

from tests.util import load_golden

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "ollipticid_amd", "Olliptic.x")


def test_cli_cfg1_writes_reference_files(tmp_path):
    meta, _ = load_golden("cfg1_tol")
    print("cli cfg1", flush=True)
    out = subprocess.run([EXE] + meta["argv"] + ["-o", "run", "-p3d", "yes"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    run = tmp_path / "run"
    for f in ("run.run", "stdout0.out", "Loo.gp", "u-l4b0-0.xyz"):
        assert (run / f).exists(), f
    hist = np.loadtxt(run / "Loo.gp")
    levels = meta["levels"]
    assert hist.shape[0] % levels == 0
    cycles = hist.shape[0] // levels
    assert 5 <= cycles <= 10
    # the last top-level post-smoothing norm is below the tolerance
    assert "ADM mass[0]" in out.stdout
    adm = float(out.stdout.split("ADM mass[0] = ")[-1].split()[0])
    assert abs(adm - meta["adm_mass"][0]) < 1e-6
    xyz = np.loadtxt(run / "u-l4b0-0.xyz")
    assert xyz.shape == (33 ** 3, 4)


def test_cli_rejects_other_equations(tmp_path):
    out = subprocess.run([EXE, "-eq", "Poisson", "-B", "Robin", "-N", "33"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "not solved here" in out.stdout
