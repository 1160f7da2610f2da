"""The analytic approximation on the GPU: -m AppPuncture (ApproximateID,
elliptic.cc:1977-2222) against the reference's runs (tests/golden/app_*:
every box's u to its 12 printed digits) and against the oracle on every
point (1e-13 relative: device sqrt/log against the host's), the printed
puncture values (Get_Puncture_App, host arithmetic) equal to the oracle's
bit for bit and to the reference's printed lines character for character,
and the approximation as the V-cycles' initial guess converging to the same
solution as the reference's u = 0 start."""
import os
import subprocess

import numpy as np
import pytest

import ollipticid_amd as oa
from oracle import oracle as O
from tests.util import determined_mask, golden_u, load_golden

pytestmark = pytest.mark.gpu
REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(REPO, "ollipticid_amd", "Olliptic.x")
CASES = ["app_two", "app_origin", "app_three"]


@pytest.mark.parametrize("name", CASES)
def test_app_matches_reference_and_oracle(name):
    meta, arrays = load_golden(name)
    print("app case", name, flush=True)
    p = oa.problem_from_argv(meta["argv"])
    assert p.method == oa.METHOD_APP
    s = O.Solver.from_argv(meta["argv"])
    s.approx_id()
    with oa.Hierarchy(p) as h:
        assert h.mgrid == 0
        cycles, _, res = h.solve()
        assert cycles == 0 and res == 0.0
        st = meta["stride"]
        for ib, ref in enumerate(golden_u(meta, arrays, h.boxes())):
            u = h.get_field(ib, oa.F_U)
            o = s.field(ib, O.F_U)
            np.testing.assert_allclose(u, o, rtol=1e-13, atol=1e-300)
            np.testing.assert_allclose(u[::st, ::st, ::st], ref, rtol=1e-11, atol=1e-300)
        pu = h.puncture_approx()
        assert pu == s.puncture_approx()
        np.testing.assert_allclose(pu, meta["puncture_u"], rtol=6e-8)
    s.close()


def test_cli_app_prints_reference_lines(tmp_path):
    meta, _ = load_golden("app_three")
    out = subprocess.run([EXE] + meta["argv"] + ["-o", "run", "-p1d", "no", "-pint", "no"],
                         cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = [ln for ln in out.stdout.splitlines() if "Approximate ID" in ln or " u : " in ln]
    assert lines == meta["stdout"]


def test_engine_refuses_app_requests():
    meta, _ = load_golden("app_two")
    e = oa.Engine()
    try:
        with pytest.raises(oa.OllError):
            e.admit(oa.problem_from_argv(meta["argv"]), 0)
    finally:
        e.close()


def test_approx_initial_guess_converges_to_reference():
    """cfg1 iterated to stagnation from the approximation: every determined
    point within 1e-9 of the reference's converged u, like the u = 0 start."""
    meta, arrays = load_golden("cfg1_conv")
    out = {}
    for guess in (oa.GUESS_ZERO, oa.GUESS_APPROX):
        p = oa.problem_from_argv(meta["argv"])
        p.initial_guess = guess
        with oa.Hierarchy(p) as h:
            cycles, _, res = h.solve()
            worst = 0.0
            for b, ref, ib in zip(h.boxes(), golden_u(meta, arrays, h.boxes()), range(h.nbox)):
                m = determined_mask(b, h.mgrid)
                worst = max(worst, float(np.abs(h.get_field(ib, oa.F_U) - ref)[m].max()))
            out[guess] = (cycles, res, worst)
    print("cfg1 to stagnation: zero start %s, approximation start %s" % (out[0], out[1]), flush=True)
    for cycles, res, worst in out.values():
        assert res < 1e-14 and worst <= 1e-9
