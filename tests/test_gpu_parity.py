"""GPU parity: every HIP stage against the oracle on the same inputs, through
the C ABI.

Stage tests compare with the oracle in multi-colour order (the GPU's order;
within one colour the result is order independent), so the only differences
are floating-point contraction and pow() versus repeated multiplication:
tolerance 1e-12 relative.  Whole solves are compared with the reference's own
converged solution (tests/golden, reference iterated to stagnation) on every
point the discrete problem determines: tolerance 1e-9 absolute.
"""
import numpy as np
import pytest

import ollipticid_amd as oa
from oracle import oracle as O
from tests.util import determined_mask, golden_u, load_golden

pytestmark = pytest.mark.gpu

PUNC = ["-eq", "Punctures", "-B", "Robin", "-BA", "0", "-Bn", "1"]
# two punctures whose finer boxes merge, 4th order, spins and momenta
SMALL_O4 = PUNC + ["-NumBH", "2", "-bhmp1", "0.4", "-bhx1", "1.1", "-bhy1", "0.3",
                   "-bhz1", "-0.2", "-bhpx1", "0.05", "-bhpy1", "0.12", "-bhsz1", "0.05",
                   "-bhmp2", "0.35", "-bhx2", "-1.3", "-bhy2", "-0.4", "-bhz2", "0.5",
                   "-bhpx2", "-0.05", "-bhpz2", "-0.1", "-bhsx2", "0.04",
                   "-L", "48", "-N", "25", "-l", "3", "-O", "4", "-nu2", "4", "-s", "1e-12"]
SCALAR_O6 = ["-eq", "Punctures_scalar_field", "-B", "Robin", "-psc", "2", "-phi0", "0.02",
             "-a2", "1", "-r0", "10", "-sigma", "4", "-NumBH", "1", "-bhmp1", "0.5",
             "-bhx1", "0.7", "-bhpy1", "0.1", "-L", "64", "-N", "25", "-l", "2", "-O", "6",
             "-nu2", "4", "-s", "1e-12"]
O8 = PUNC + ["-NumBH", "1", "-bhmp1", "0.8", "-bhx1", "0.5", "-bhpy1", "0.15", "-bhsz1", "0.1",
             "-L", "50", "-N", "25", "-l", "2", "-O", "8", "-nu2", "4", "-s", "1e-12"]
# 2nd order, two punctures; the finest box's points that coincide with its
# parent's nodes have odd i + j + k, so its sweeps take colour 1 first
O2_FLIP = PUNC + ["-NumBH", "2", "-bhmp1", "0.5", "-bhx1", "0.5", "-bhpy1", "0.1",
                  "-bhmp2", "0.4", "-bhx2", "-3", "-bhy2", "2", "-bhsz2", "0.1",
                  "-L", "32", "-N", "25", "-l", "4", "-O", "2", "-nu2", "4", "-s", "1e-12"]
CASES = {"o4_merge": SMALL_O4, "sf_o6": SCALAR_O6, "o8": O8, "o2_flip": O2_FLIP}


def pair(argv):
    print("case argv:", " ".join(argv), flush=True)
    h = oa.Hierarchy(oa.problem_from_argv(argv))
    s = O.Solver.from_argv(argv, sweep_order=O.COLOUR)
    assert [b["n"] for b in h.boxes()] == [b["n"] for b in s.boxes()]
    return h, s


def seed_state(h, s, seed):
    """Same smooth-plus-noise u on both (c + u stays positive)."""
    rng = np.random.default_rng(seed)
    for ib, b in enumerate(s.boxes()):
        n = b["n"]
        z, y, x = np.meshgrid(*[np.linspace(0, 1, n[a]) for a in (2, 1, 0)], indexing="ij")
        u = 0.01 * np.sin(3 * x + 1) * np.cos(2 * y) * (1 + z) + 1e-3 * rng.standard_normal(x.shape)
        s.field(ib, O.F_U)[:] = u
        h.set_field(ib, oa.F_U, u)


def rel_err(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("case", list(CASES))
def test_coefficients(case):
    h, s = pair(CASES[case])
    fields = [(oa.F_B, O.F_B), (oa.F_C, O.F_C), (oa.F_RHO, O.F_RHO), (oa.F_U, O.F_U)]
    if "scalar" in " ".join(CASES[case]):
        fields += [(oa.F_D, O.F_D), (oa.F_E, O.F_E)]
    for ib in range(s.nbox):
        for fg, fo in fields:
            g, o = h.get_field(ib, fg), s.field(ib, fo)
            # c is ~1e15 at a puncture on a grid point; compare relative per point
            den = np.maximum(np.abs(o), 1e-300)
            assert (np.abs(g - o) / den)[np.abs(o) > 0].max(initial=0) < 1e-12, (case, ib, fg)
            assert np.all((o == 0) <= (np.abs(g) < 1e-300)), (case, ib, fg)


@pytest.mark.parametrize("case", list(CASES))
def test_residual_every_level(case):
    h, s = pair(CASES[case])
    seed_state(h, s, 1)
    for lv in range(s.levels):
        ng = h.residual(lv)
        no = max(s.residual(ib) for ib in s.level_boxes(lv))
        assert abs(ng - no) <= 1e-12 * no, (case, lv)
        for ib in s.level_boxes(lv):
            R = s.box_info(ib)["R"]
            g = h.get_field(ib, oa.F_R)[R:-R, R:-R, R:-R]
            o = s.field(ib, O.F_R)[R:-R, R:-R, R:-R]
            assert rel_err(g, o) < 1e-12, (case, lv, ib)


@pytest.mark.parametrize("case", list(CASES))
def test_smooth_sweeps_every_level(case):
    h, s = pair(CASES[case])
    seed_state(h, s, 2)
    for lv in range(s.levels):
        ng = h.smooth(lv, 2)
        no = max(s.smooth(ib, 2) for ib in s.level_boxes(lv))
        assert abs(ng - no) <= 1e-11 * no, (case, lv, ng, no)
        for ib in s.level_boxes(lv):
            assert rel_err(h.get_field(ib, oa.F_U), s.field(ib, O.F_U)) < 1e-12, (case, lv, ib)


@pytest.mark.parametrize("case", ["o4_merge", "sf_o6"])
def test_caller_rho_reaches_smoother_and_defect(case):
    """The smoother and the defect skip the rho load where only zeros can
    be (points no restriction writes, the whole top level): a rho the caller
    writes, with oll_set_field or oll_transfer, must still be read at every
    interior point of every level."""
    h, s = pair(CASES[case])
    try:
        seed_state(h, s, 5)
        rng = np.random.default_rng(5)
        top = s.levels - 1
        for lv in range(s.levels):
            for ib in s.level_boxes(lv):
                if lv == top - 1:  # this level's rho comes by a transfer below
                    continue
                # interior points only: the ghost shell's rho is the Robin
                # constant the boundary stages take
                R = s.box_info(ib)["R"]
                rho = s.field(ib, O.F_RHO).copy()
                inner = rho[R:-R, R:-R, R:-R]
                inner[:] = 1e-3 * rng.standard_normal(inner.shape)
                s.field(ib, O.F_RHO)[:] = rho
                h.set_field(ib, oa.F_RHO, rho)
        # interior rho of the level below the top: u of the top level,
        # Lagrange order 3
        h.transfer(top, oa.F_U, top - 1, oa.F_RHO, 1, 1, 3, True, False)
        s.transfer(top, O.F_U, top - 1, O.F_RHO, 1, 1, 3, 1, 0)
        for lv in range(s.levels):
            ng = h.smooth(lv, 2)
            no = max(s.smooth(ib, 2) for ib in s.level_boxes(lv))
            assert abs(ng - no) <= 1e-11 * no, (case, lv, ng, no)
            for ib in s.level_boxes(lv):
                assert rel_err(h.get_field(ib, oa.F_U), s.field(ib, O.F_U)) < 1e-12, (case, lv, ib)
        # one FAS step from the top: its defect rho - N(u) of the top level
        # (rho the caller's) is restricted into every coarser level's source
        ng, no = h.fas(top), s.fasn(top)
        assert abs(ng - no) <= 1e-9 * no, (case, ng, no)
        for ib, b in enumerate(s.boxes()):
            g, o = h.get_field(ib, oa.F_U), s.field(ib, O.F_U)
            assert np.abs(g - o).max() <= 1e-11 * max(1.0, np.abs(o).max()), (case, ib)
    finally:
        h.close()
        s.close()


@pytest.mark.parametrize("case", list(CASES))
def test_robin_and_boundary_operator(case):
    """Robin operator on a seeded state (non-zero everywhere on the shell),
    then the Robin update, which leaves operator == rho on the shell."""
    h, s = pair(CASES[case])
    seed_state(h, s, 3)
    for lv in range(s.mgrid + 1):
        ib = s.level_boxes(lv)[0]
        R = s.box_info(ib)["R"]
        shell = np.ones(s.field(ib, O.F_U).shape, dtype=bool)
        shell[R:-R, R:-R, R:-R] = False
        h.boundary_operator(lv)
        s.boundary_operator(ib)
        g, o = h.get_field(ib, oa.F_R)[shell], s.field(ib, O.F_R)[shell]
        assert rel_err(g, o) < 1e-12, (case, lv)
        h.robin(lv)
        s.robin(ib)
        assert rel_err(h.get_field(ib, oa.F_U), s.field(ib, O.F_U)) < 1e-12, (case, lv)
        # after the update the Robin equation holds on the shell (rho = 0)
        h.boundary_operator(lv)
        assert np.abs(h.get_field(ib, oa.F_R)[shell]).max() < 1e-12 * np.abs(o).max()


@pytest.mark.parametrize("case", list(CASES))
def test_fas_cycles_match_oracle(case):
    """Three whole FAS V-cycles from u = 0: every field value of every box."""
    h, s = pair(CASES[case])
    top = s.levels - 1
    for cyc in range(3):
        ng = h.fas(top)
        no = s.fasn(top)
        assert abs(ng - no) <= 1e-9 * no, (case, cyc, ng, no)
    for ib, b in enumerate(s.boxes()):
        g, o = h.get_field(ib, oa.F_U), s.field(ib, O.F_U)
        assert np.abs(g - o).max() <= 1e-11 * max(1.0, np.abs(o).max()), (case, ib)


TRANSFERS = [
    # (from, to) relative to the top level, region, method, order, all, to_zero
    ("restrict u", 0, -1, 1, 1, 3, False, False),
    ("restrict w", 0, -1, 0, 1, 2, True, False),
    ("prolong", -1, 0, 0, 0, 1, True, True),
    ("ghost fill", -1, 0, 2, 1, 5, True, False),
    ("two levels up", -2, 0, 1, 1, 4, False, True),
]


@pytest.mark.parametrize("case", ["o4_merge", "o2_flip"])
def test_transfer_matches_oracle(case):
    """oll_transfer (multigrid::Transferir) against the oracle's transfer
    for the restriction, prolongation and ghost-fill patterns of the
    V-cycle and one more: every point of every target box, from a seeded
    u into a v that starts equal on both sides."""
    h, s = pair(CASES[case])
    try:
        seed_state(h, s, 7)
        top = s.levels - 1
        for name, df, dt, region, method, order, all_, to_zero in TRANSFERS:
            fl, tl = top + df, top + dt
            if min(fl, tl) < 0:
                continue
            print("transfer", case, name, flush=True)
            for ib in s.level_boxes(tl):
                shape = s.field(ib, O.F_V).shape
                v0 = 0.25 + 0.01 * (np.arange(np.prod(shape)) % 7).reshape(shape)
                s.field(ib, O.F_V)[:] = v0
                h.set_field(ib, oa.F_V, v0)
            h.transfer(fl, oa.F_U, tl, oa.F_V, region, method, order, all_, to_zero)
            s.transfer(fl, O.F_U, tl, O.F_V, region, method, order, int(all_), int(to_zero))
            for ib in s.level_boxes(tl):
                g, o = h.get_field(ib, oa.F_V), s.field(ib, O.F_V)
                assert rel_err(g, o) < 1e-12, (case, name, ib, rel_err(g, o))
    finally:
        h.close()
        s.close()


def test_colour_order_matches_oracle():
    """The GPU's per-box colour order is the oracle's, and the 2nd-order case
    has a box swept colour 1 first."""
    h, s = pair(O2_FLIP)
    got = [b["colour_order"] for b in h.boxes()]
    assert got == [s.colour_flip(ib) for ib in range(s.nbox)]
    assert 1 in got


# the linear Brill-waves equation (Lap u + Lq u = 0), 2nd order, refined
# around x = 0.5: two boxes are swept colour 1 first
BRILL_FLIP = ["-eq", "Brill-waves", "-B", "Robin", "-NumBH", "1", "-bhx1", "0.5", "-L", "20",
              "-N", "25", "-l", "4", "-O", "2", "-nu2", "4", "-s", "1e-12"]


def test_mixed_smooth_two_colour_order():
    """Mixed-precision (FP32 correction) sweeps on every level of a
    2nd-order case with descending-order boxes, against the oracle's FP64
    sweeps in the same colour order.  The equation is linear, so the FP32
    sweeps on the correction equation are the FP64 sweeps up to FP32
    rounding: the change of u agrees to 1e-5 relative; a wrong colour order
    changes it at O(1)."""
    p = oa.problem_from_argv(BRILL_FLIP)
    p.precision = oa.PREC_MIXED
    h = oa.Hierarchy(p)
    s = O.Solver.from_argv(BRILL_FLIP, sweep_order=O.COLOUR)
    assert sum(b["colour_order"] for b in h.boxes()) >= 2
    seed_state(h, s, 4)
    for lv in range(1, s.levels):
        u0 = {ib: s.field(ib, O.F_U).copy() for ib in s.level_boxes(lv)}
        h.smooth(lv, 2)
        for ib in s.level_boxes(lv):
            s.smooth(ib, 2)
            dg = h.get_field(ib, oa.F_U) - u0[ib]
            do = s.field(ib, O.F_U) - u0[ib]
            print("level", lv, "box", ib, "colour order", s.colour_flip(ib),
                  "rel diff %.2e" % rel_err(dg, do), flush=True)
            assert rel_err(dg, do) < 1e-5, (lv, ib, s.colour_flip(ib), rel_err(dg, do))


def test_solve_cfg1_matches_reference():
    """cfg1 to stagnation: every determined point within 1e-9 of the
    reference's converged u (tests/golden/cfg1_conv)."""
    meta, arrays = load_golden("cfg1_conv")
    h = oa.Hierarchy(oa.problem_from_argv(meta["argv"]))
    cycles, hist, res = h.solve()
    print("cfg1 conv: cycles", cycles, "res", res, "ms", h.last_solve_ms(), flush=True)
    assert res < 1e-14
    worst = 0.0
    for ib, (b, ref) in enumerate(zip(h.boxes(), golden_u(meta, arrays, h.boxes()))):
        m = determined_mask(b, h.mgrid)
        worst = max(worst, np.abs(h.get_field(ib, oa.F_U) - ref)[m].max())
    print("cfg1 max |u - u_ref|:", worst, flush=True)
    assert worst <= 1e-9
    adm, _ = h.adm_mass()
    np.testing.assert_allclose(adm, meta["adm_mass"], rtol=5.01e-9)


def test_solve_small_converging_case_matches_reference():
    """Four punctures, merged boxes, 2nd order: the reference converges to
    3.6e-15; every determined point within 1e-9 of its u."""
    meta, arrays = load_golden("small_o2_4bh")
    h = oa.Hierarchy(oa.problem_from_argv(meta["argv"]))
    cycles, hist, res = h.solve()
    print("small_o2_4bh: cycles", cycles, "res", res, flush=True)
    worst = 0.0
    for ib, (b, ref) in enumerate(zip(h.boxes(), golden_u(meta, arrays, h.boxes()))):
        m = determined_mask(b, h.mgrid)
        worst = max(worst, np.abs(h.get_field(ib, oa.F_U) - ref)[m].max())
    assert worst <= 1e-9


def test_adm_mass_matches_oracle():
    h, s = pair(SMALL_O4)
    top = s.levels - 1
    for _ in range(2):
        h.fas(top)
        s.fasn(top)
    pg, sg = h.adm_mass()
    po, so = s.adm_mass()
    np.testing.assert_allclose(pg, po, rtol=1e-12)
    np.testing.assert_allclose(sg, so, rtol=1e-10)
