"""The oracle (plain-C restatement, oracle/) against the reference's own
outputs (tests/golden, written by tests/golden/make_golden.py from the
reference binary compiled from its sources).  In the reference's
lexicographic order the oracle must reproduce the residual history
(Loo.gp, 18 digits) exactly, the cycle count, the converged u to the
fixtures' 12 printed digits, the puncture ADM masses and every box's
surface ADM mass (elliptic.cc:2471-2606, "AMDmass=") to their 9 printed
digits."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.util import golden_u, load_golden

FAST = ["cfg1_tol", "small_o4_offgrid", "small_o6_merge", "small_sf_3bh", "small_o2_4bh",
        "small_poisson_o2", "small_poisson_o6", "small_brill_o4",
        "small_em_o2", "small_em_o4", "small_em_o6", "small_em_o8"]


@pytest.mark.parametrize("name", FAST)
def test_oracle_matches_reference(name):
    meta, arrays = load_golden(name)
    s = O.Solver.from_argv(meta["argv"])
    boxes = s.boxes()
    assert [(b["level"], b["n"]) for b in boxes] == [(b["level"], b["n"]) for b in meta["boxes"]]
    cycles, hist, res = s.solve()
    assert cycles == meta["cycles"]
    ref_hist = np.array(meta["history"])
    assert hist.shape == ref_hist.shape
    np.testing.assert_array_equal(hist[:, 0], ref_hist[:, 0])
    # Loo.gp prints 18 significant digits
    np.testing.assert_allclose(hist[:, 1], ref_hist[:, 1], rtol=1e-15, atol=0)
    st = meta["stride"]
    for ib, ref in enumerate(golden_u(meta, arrays, boxes)):
        u = s.field(ib, O.F_U)[::st, ::st, ::st]
        scale = max(1.0, np.abs(ref).max())
        assert np.abs(u - ref).max() <= 1e-11 * scale + 1e-14
    check_adm(s, meta)


def check_adm(s, meta):
    """Puncture and per-box surface ADM masses against the reference's
    stdout (9 significant digits, so half a unit of the 9th digit)."""
    if not meta["adm_mass"]:  # printed for the puncture equations only
        return
    adm, surf = s.adm_mass()
    np.testing.assert_allclose(adm, meta["adm_mass"], rtol=5.01e-9)
    ref_surf = [b["surface_adm"] for b in meta["boxes"]]
    assert len(ref_surf) == len(surf) and all(v is not None for v in ref_surf)
    for ib, (got, ref) in enumerate(zip(surf, ref_surf)):
        assert abs(got - ref) <= 5.01e-9 * abs(ref) + 1e-300, (ib, got, ref)


@pytest.mark.slow
def test_oracle_matches_reference_cfg2():
    """The equal-mass binary (N=65, 9 levels, 4th order, tol 1e-11): 17
    cycles, Loo.gp history bit for bit, u to print precision and the ADM
    masses; about 90 s on one core."""
    test_oracle_matches_reference("cfg2_tol")


def test_oracle_converged_cfg1_field():
    """cfg1 iterated to stagnation (-s 1e-16): 30 cycles, u to print precision."""
    meta, arrays = load_golden("cfg1_conv")
    s = O.Solver.from_argv(meta["argv"])
    cycles, hist, res = s.solve()
    assert cycles == meta["cycles"] == 30
    for ib, ref in enumerate(golden_u(meta, arrays, s.boxes())):
        assert np.abs(s.field(ib, O.F_U) - ref).max() < 1e-14


def test_continued_binary_fixture():
    """tests/golden/cfg2_continued: the reference's binary run (N=65, 9
    levels, cfg2_allconv) replayed by the oracle to 1e-14 and continued 150
    V-cycles until no stored value moves (< 1e-14 over the last 20).  It
    agrees with the reference's own stopped run to 1e-10 on every level but
    the base box, which the reference's stopping rule leaves about 2e-9
    short of convergence."""
    cont, carr = load_golden("cfg2_continued")
    ref, rarr = load_golden("cfg2_allconv")
    assert cont["reference_cycles"] == ref["cycles"] == 73
    assert cont["replay_max_diff"] <= 1e-14
    assert cont["moved_over_last_20_cycles"] <= 1e-14
    per_level = {}
    for b in cont["boxes"]:
        key = "u_l%d_b%d" % (b["level"], b["box"])
        d = float(np.abs(carr[key] - rarr[key]).max())
        per_level[b["level"]] = max(per_level.get(b["level"], 0.0), d)
    mgrid = 2
    assert all(d <= 1e-10 for lv, d in per_level.items() if lv > mgrid), per_level
    assert 1e-9 < per_level[mgrid] < 5e-9, per_level
