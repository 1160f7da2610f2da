"""The oracle's analytic approximation (ApproximateID, elliptic.cc:1977-2222;
Get_Puncture_App, :2232-2444) against the reference's -m AppPuncture runs
(tests/golden/app_*: u of every box to 12 printed digits, the puncture
values to 8) and against what the formula fixes on its own: a puncture with
neither momentum nor spin contributes nothing, the spin term alone is
symmetric under x -> -x about the puncture, and the terms of several
punctures add."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.util import load_golden

CASES = ["app_two", "app_origin", "app_three"]


@pytest.mark.parametrize("name", CASES)
def test_matches_reference_fields(name):
    meta, arrays = load_golden(name)
    s = O.Solver.from_argv(meta["argv"])
    try:
        assert s.mgrid == 0
        assert [(b["level"], b["n"]) for b in s.boxes()] == [(b["level"], b["n"]) for b in meta["boxes"]]
        s.approx_id()
        st = meta["stride"]
        seen = {}
        for ib, b in enumerate(s.boxes()):
            lv = b["level"]
            k = seen.get(lv, 0)
            seen[lv] = k + 1
            ref = arrays["u_l%d_b%d" % (lv, k)]
            u = s.field(ib, O.F_U)[::st, ::st, ::st]
            # 12 significant digits printed
            np.testing.assert_allclose(u, ref, rtol=6e-12, atol=1e-300)
        np.testing.assert_allclose(s.puncture_approx(), meta["puncture_u"], rtol=6e-8)
    finally:
        s.close()


PUNC = ["-eq", "Punctures", "-B", "Robin", "-NumBH", "1", "-L", "20", "-N", "25", "-l", "2",
        "-O", "4", "-m", "AppPuncture"]


def field_of(argv):
    s = O.Solver.from_argv(argv)
    s.approx_id()
    u = s.field(s.nbox - 1, O.F_U).copy()
    s.close()
    return u


def test_bare_puncture_gives_zero():
    assert np.all(field_of(PUNC + ["-bhmp1", "0.7", "-bhx1", "0.3"]) == 0.0)


def test_spin_term_is_even_about_the_puncture():
    # puncture at the origin on a box symmetric about it: u(-x) = u(x)
    u = field_of(PUNC + ["-bhmp1", "0.7", "-bhsz1", "0.2"])
    np.testing.assert_allclose(u, u[:, :, ::-1], rtol=1e-13, atol=0)
    assert np.abs(u).max() > 0


def base_field(argv):
    s = O.Solver.from_argv(argv)
    s.approx_id()
    u = s.field(0, O.F_U).copy()
    s.close()
    return u


def test_terms_of_two_punctures_add():
    # the base box is the same grid whatever the punctures
    a = ["-bhmp1", "0.5", "-bhx1", "1.0", "-bhpy1", "0.2", "-bhsz1", "0.1"]
    b = ["-bhmp2", "0.4", "-bhx2", "-1.0", "-bhpx2", "0.1", "-bhsy2", "-0.05"]
    b_alone = [t.replace("2", "1") if t.startswith("-bh") else t for t in b]
    two = PUNC[:4] + ["-NumBH", "2"] + PUNC[6:] + a + b
    np.testing.assert_allclose(base_field(two), base_field(PUNC + a) + base_field(PUNC + b_alone),
                               rtol=1e-14, atol=1e-300)
