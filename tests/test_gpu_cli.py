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
    out = subprocess.run([EXE, "-eq", "Bowen-York", "-B", "Robin", "-N", "33"], cwd=tmp_path,
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "not solved here" in out.stdout


def test_cli_poisson_solves_without_adm_lines(tmp_path):
    """-eq Poisson: no ADM mass lines (elliptic.cc:1939-1941 prints them for
    the puncture equations only) and the finest box's u within 1e-9 of the
    reference's converged run."""
    meta, arrays = load_golden("small_poisson_o2")
    print("cli poisson", flush=True)
    out = subprocess.run([EXE] + meta["argv"] + ["-o", "run", "-p3d", "yes", "-p1d", "no",
                                                  "-pint", "no"],
                         cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "ADM mass" not in out.stdout and "AMDmass" not in out.stdout
    top = meta["levels"] - 1
    xyz = np.loadtxt(tmp_path / "run" / ("u-l%db0-0.xyz" % top))
    ref = arrays["u_l%d_b0" % top]
    assert xyz.shape[0] == ref.size
    assert np.abs(xyz[:, 3] - ref.ravel()).max() <= 1e-9


def test_cli_punctures_em_prints_masses_and_charges(tmp_path):
    """-eq Punctures_EM (order 2): the reference's closing mp/qp lines
    (elliptic.cc:1333-1339), no ADM lines, finest u within 1e-9 of the
    reference's converged run."""
    meta, arrays = load_golden("small_em_o2")
    print("cli punctures_em", flush=True)
    out = subprocess.run([EXE] + meta["argv"] + ["-o", "run", "-p3d", "yes", "-p1d", "no",
                                                  "-pint", "no"],
                         cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = [ln for ln in out.stdout.splitlines() if ln.startswith(("mp [", "qp ["))]
    assert lines == ["mp [0] = 5.00000000e-01", "mp [1] = 4.00000000e-01",
                     "qp [0] = 2.00000000e-01", "qp [1] = -1.00000000e-01"]
    assert "ADM mass" not in out.stdout
    top = meta["levels"] - 1
    xyz = np.loadtxt(tmp_path / "run" / ("u-l%db0-0.xyz" % top))
    assert np.abs(xyz[:, 3] - arrays["u_l%d_b0" % top].ravel()).max() <= 1e-9
