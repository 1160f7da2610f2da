"""Poisson, Brill-waves and Punctures_EM on the GPU through the same kernels
as the puncture equations (the equation only changes the terms of N(u) next
to the Laplacian and the set-up of the coefficients, rho and u).

Reference problems (elliptic.h:357-398, 510-626, operator.cc:640-725,
1010-1100): Poisson Lap u + u^2 = Gaussian_source with the Gaussian as
initial and boundary data, Robin value 0; Brill-waves Lap u + Lq u = 0 with
Lq the Brill seed, u = 1 initially, Robin value 1; Punctures_EM with the
charged punctures' coefficients, u = 1, Robin value 1, at orders 4-8 swept
in 5 colours (its one-sided first derivatives reach 4 points).  Stage
tests against the oracle in the GPU's multi-colour order (1e-12 relative);
whole solves iterated to stagnation against the reference's own runs
(tests/golden/small_poisson_*, small_brill_*, small_em_o2) on every
determined point (1e-9 absolute)."""
import numpy as np
import pytest

import ollipticid_amd as oa
from oracle import oracle as O
from tests.util import determined_mask, golden_u, load_golden

pytestmark = pytest.mark.gpu
CASES = ["small_poisson_o2", "small_poisson_o6", "small_brill_o4", "small_brill_o8", "small_em_o2"]
# Punctures_EM at orders 4, 6 and 8 (5-colour sweeps, one-sided derivatives):
# stage parity only, the reference's own iteration stagnating near 1e-6
EM_STAGES = ["small_em_o4", "small_em_o6", "small_em_o8"]


def pair(name):
    meta, _ = load_golden(name)
    print("equation case", name, " ".join(meta["argv"]), flush=True)
    h = oa.Hierarchy(oa.problem_from_argv(meta["argv"]))
    s = O.Solver.from_argv(meta["argv"], sweep_order=O.COLOUR)
    assert [b["n"] for b in h.boxes()] == [b["n"] for b in s.boxes()]
    return h, s


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def em_coefficients(h, ib):
    """The thirteen Punctures_EM coefficients (elliptic.h:1458-1530) from the
    potentials and gradients the GPU holds, in the oracle's field order."""
    eta, phi = h.get_field(ib, oa.F_B), h.get_field(ib, oa.F_D)
    ge = [h.get_field(ib, f) for f in (oa.F_ETA_X, oa.F_ETA_Y, oa.F_ETA_Z)]
    gf = [h.get_field(ib, f) for f in (oa.F_PHI_X, oa.F_PHI_Y, oa.F_PHI_Z)]
    f2 = sum(g * g for g in gf)
    fe = sum(a * b for a, b in zip(gf, ge))
    e2 = sum(g * g for g in ge)
    return ([3 * eta, 3 * eta ** 2 - 0.25 * phi ** 2, eta ** 3 - 0.25 * eta * phi ** 2,
             -0.25 * phi ** 2] + [0.5 * phi * g for g in gf] +
            [0.5 * phi * (eta * a - phi * b) for a, b in zip(gf, ge)] +
            [-0.25 * f2, 0.5 * phi * fe - 0.5 * eta * f2,
             -0.25 * phi ** 2 * e2 + 0.5 * phi * eta * fe + 0.25 * (1 - eta ** 2) * f2])


@pytest.mark.parametrize("name", CASES + EM_STAGES)
def test_setup_matches_oracle(name):
    h, s = pair(name)
    em = "_em_" in name
    for ib in range(s.nbox):
        pairs = [(oa.F_RHO, O.F_RHO), (oa.F_U, O.F_U)]
        if not em:
            pairs += [(oa.F_B, O.F_B), (oa.F_C, O.F_C)]
        for fg, fo in pairs:
            g, o = h.get_field(ib, fg), s.field(ib, fo)
            assert np.abs(g - o).max() <= 1e-13 * max(np.abs(o).max(), 1.0), (name, ib, fg)
        if em:
            # relative to each coefficient's largest value: the two
            # punctures' terms cancel to round-off in places
            for c, (g, fo) in enumerate(zip(em_coefficients(h, ib), O.F_EM)):
                o = s.field(ib, fo)
                assert np.abs(g - o).max() <= 1e-13 * np.abs(o).max(), (name, ib, c)
    h.close()
    s.close()


@pytest.mark.parametrize("name", CASES + EM_STAGES)
def test_stages_match_oracle(name):
    """Residual and two sweeps on every level, then three whole FAS cycles."""
    h, s = pair(name)
    for lv in range(s.levels):
        ng = h.residual(lv)
        no = max(s.residual(ib) for ib in s.level_boxes(lv))
        assert abs(ng - no) <= 1e-12 * max(no, 1e-300), (name, lv, ng, no)
        ng = h.smooth(lv, 2)
        no = max(s.smooth(ib, 2) for ib in s.level_boxes(lv))
        assert abs(ng - no) <= 1e-11 * max(no, 1e-300), (name, lv, ng, no)
        for ib in s.level_boxes(lv):
            assert rel_err(h.get_field(ib, oa.F_U), s.field(ib, O.F_U)) < 1e-12, (name, lv, ib)
    top = s.levels - 1
    for cyc in range(3):
        ng = h.fas(top)
        no = s.fasn(top)
        assert abs(ng - no) <= 1e-9 * no, (name, cyc, ng, no)
    for ib in range(s.nbox):
        g, o = h.get_field(ib, oa.F_U), s.field(ib, O.F_U)
        assert np.abs(g - o).max() <= 1e-11 * max(1.0, np.abs(o).max()), (name, ib)
    h.close()
    s.close()


def solve_to_stagnation(argv, precision):
    p = oa.problem_from_argv(argv)
    p.precision = precision
    p.stop_rule = oa.STOP_TOLERANCE
    p.tol = 1e-30
    p.patience = 8
    p.max_cycles = 400
    h = oa.Hierarchy(p)
    cycles, _, res = h.solve()
    return h, cycles, res


@pytest.mark.parametrize("precision", [oa.PREC_FP64, oa.PREC_MIXED])
@pytest.mark.parametrize("name", CASES)
def test_solution_matches_reference(name, precision):
    if "_em_" in name and precision == oa.PREC_MIXED:
        pytest.skip("Punctures_EM runs in FP64 only")
    meta, arrays = load_golden(name)
    h, cycles, res = solve_to_stagnation(meta["argv"], precision)
    worst = 0.0
    for ib, (b, ref) in enumerate(zip(h.boxes(), golden_u(meta, arrays, h.boxes()))):
        m = determined_mask(b, h.mgrid)
        worst = max(worst, float(np.abs(h.get_field(ib, oa.F_U) - ref)[m].max()))
    print(name, "precision", precision, "cycles", cycles, "res", res, "max |u - u_ref|", worst,
          flush=True)
    h.close()
    assert res < 1e-13 and worst <= 1e-9


def test_engine_batches_equations_together():
    """The equations' problems admitted together to one engine (the 3- and
    5-colour 4th-order groups side by side): each result equals the same
    problem solved alone, bit for bit."""
    probs = {}
    for t, name in enumerate(CASES + ["small_em_o4"]):
        meta, _ = load_golden(name)
        p = oa.problem_from_argv(meta["argv"])
        p.stop_rule = oa.STOP_TOLERANCE
        p.tol = 1e-30
        p.patience = 4
        p.max_cycles = 20
        probs[t] = p
    alone = {}
    for t, p in probs.items():
        with oa.Hierarchy(p) as h:
            cycles, _, res = h.solve()
            alone[t] = (cycles, res, [h.get_field(ib, oa.F_U) for ib in range(h.nbox)])
    e = oa.Engine()
    try:
        for t, p in probs.items():
            assert e.admit(p, t)
        got = {}
        for _ in range(400):
            active = e.step()
            for tag, cycles, res, h in e.pop_hier(probs):
                got[tag] = (cycles, res, [h.get_field(ib, oa.F_U) for ib in range(h.nbox)])
                h.close()
            if active == 0:
                break
        assert sorted(got) == sorted(probs)
        for t in probs:
            assert got[t][:2] == alone[t][:2], (t, got[t][:2], alone[t][:2])
            for ib, (g, a) in enumerate(zip(got[t][2], alone[t][2])):
                assert np.array_equal(g, a), (t, ib)
    finally:
        e.close()
