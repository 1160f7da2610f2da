"""The class API of include/ is a drop-in for the reference's on the solve
path: the reference's own src/main.cc compiles and links against include/
and the command-line layer (checked where the reference tree is present),
and tests/dropin_main.cpp, a host program using the reference's names and
signatures (interface, elliptic and its Solve, Vcycle, Set_Boundary,
App_Bound, OPERATOR, duOPERATOR; multigrid::Get_ffunction; the ffunction
iterators), builds against them; on the GPU it solves, evaluates the
operator at the solution and solves again from u = 0 to the same result,
then moves polynomials between levels with multigrid::Transferir."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(REPO, "include")
PKG = os.path.join(REPO, "ollipticid_amd")
CLI = [os.path.join(PKG, "csrc", "cli", f) for f in ("interface.cpp", "elliptic.cpp")]
REF_MAIN = "/root/reference/src/main.cc"
DROPIN = os.path.join(REPO, "tests", "_build", "dropin_main.x")
SMOKE = ["-eq", "Punctures", "-B", "Robin", "-BA", "0", "-Bn", "1", "-NumBH", "2",
         "-bhmp1", "0.5", "-bhx1", "1.5", "-bhpy1", "0.1", "-bhmp2", "0.5", "-bhx2", "-1.5",
         "-bhpy2", "-0.1", "-L", "40", "-N", "25", "-l", "3", "-O", "4", "-nu2", "6",
         "-s", "1e-10"]


@pytest.mark.skipif(not os.path.exists(REF_MAIN), reason="reference tree not present")
def test_reference_main_builds_against_include(tmp_path):
    exe = str(tmp_path / "ref_main.x")
    subprocess.check_call(["g++", "-std=c++17", "-w", "-I" + INC, REF_MAIN] + CLI +
                          ["-o", exe, "-L" + PKG, "-lolliptic_amd", "-Wl,-rpath," + PKG])
    assert os.path.getsize(exe) > 0


def test_dropin_program_built():
    assert os.path.exists(DROPIN), "run __graft_entry__.build() first"


@pytest.mark.gpu
def test_dropin_program_runs(tmp_path):
    print("case: dropin_main", " ".join(SMOKE), flush=True)
    out = subprocess.run([DROPIN] + SMOKE + ["-o", "run", "-p1d", "no", "-pint", "no"],
                         cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    lines = {ln.split()[1]: ln.split()[2:] for ln in out.stdout.splitlines()
             if ln.startswith("dropin:")}
    cycles, res = int(lines["solve"][0]), float(lines["solve"][1])
    assert 0 < cycles < 40 and res <= 1e-10
    points, worst, dn_min, dn_max = lines["operator"]
    # rho = 0 on the interior: |N(u)| is the residual of the last sweep's
    # update, below the tolerance at the solution
    assert int(points) > 0 and float(worst) <= 1e-9
    # dN/du = the Laplacian's diagonal - 7 b (c+u)^-8 < 0
    assert float(dn_max) < 0 and float(dn_min) < float(dn_max)
    # Set_Boundary restores u = 0: the second solve is the first, bit for bit
    assert lines["again"] == lines["solve"]
    # Transferir: restriction (Lagrange 3) of a cubic onto the coarser
    # level's covered interior points, prolongation (linear, to_zero) of a
    # multilinear function onto every point of the finest level
    moved, err3, pts1, err1 = lines["transfer"]
    assert int(moved) > 0 and float(err3) < 1e-11
    assert int(pts1) > 0 and float(err1) < 1e-12
    assert (tmp_path / "run" / "Loo.gp").exists()
