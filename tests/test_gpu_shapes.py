"""Launch shapes change speed only: the colour passes with 64 and with 512
threads per workgroup give bit for bit the solve of the 256-thread default
(FP64 and mixed precision, 4th and 2nd order), and the ghost-shell fill by
columns (6th and 8th order) and the defect filter by runs of points give
bit for bit the point-by-point kernels (merged and off-lattice boxes, the
scalar-field equation, 2nd to 8th order)."""
import numpy as np
import pytest

import ollipticid_amd as oa
from ollipticid_amd import native
from tests.util import load_golden

pytestmark = pytest.mark.gpu


def solve_fields(argv, precision, cycles=3):
    p = oa.problem_from_argv(argv)
    p.precision = precision
    p.max_cycles = cycles
    with oa.Hierarchy(p) as h:
        c, hist, res = h.solve()
        return c, res, [h.get_field(ib, oa.F_U) for ib in range(h.nbox)]


@pytest.mark.parametrize("precision", [oa.PREC_FP64, oa.PREC_MIXED])
@pytest.mark.parametrize("name", ["small_o2_4bh", "cfg1_tol"])
def test_sweep_threads_do_not_change_results(name, precision):
    meta, _ = load_golden(name)
    order = oa.problem_from_argv(meta["argv"]).order
    print("shapes", name, precision, flush=True)
    before = oa.sweep_threads(precision, order)
    try:
        runs = {}
        for threads in (256, 64, 512):
            oa.set_sweep_threads(precision, order, threads)
            runs[threads] = solve_fields(meta["argv"], precision)
        ref = runs[256]
        for threads in (64, 512):
            c, res, us = runs[threads]
            assert (c, res) == ref[:2], (threads, c, res, ref[:2])
            for ib, (a, b) in enumerate(zip(us, ref[2])):
                assert np.array_equal(a, b), (threads, ib)
    finally:
        oa.set_sweep_threads(precision, order, before)


@pytest.mark.parametrize("name", ["small_o2_4bh", "small_o4_offgrid", "small_o6_merge", "small_o8",
                                  "small_sf_3bh", "cfg3r_o8", "cfg5r_o6"])
def test_shared_loads_equal_points(name):
    meta, _ = load_golden(name)
    print("ghost fill", name, flush=True)
    try:
        native.set_shared_loads(False)
        ref = solve_fields(meta["argv"], oa.PREC_FP64, cycles=4)
        native.set_shared_loads(True)
        got = solve_fields(meta["argv"], oa.PREC_FP64, cycles=4)
    finally:
        native.set_shared_loads(True)
    assert got[:2] == ref[:2], (got[:2], ref[:2])
    for ib, (a, b) in enumerate(zip(got[2], ref[2])):
        assert np.array_equal(a, b), (name, ib, float(np.abs(a - b).max()))


LONG_ROWS = ["-eq", "Punctures", "-B", "Robin", "-BA", "0", "-Bn", "1", "-NumBH", "1",
             "-bhmp1", "0.5", "-bhx1", "0.3", "-bhpy1", "0.1", "-L", "64", "-N", "513", "-l", "1",
             "-O", "2", "-nu2", "2", "-s", "1e-10"]


def test_long_rows_match_oracle():
    """A 513^3 2nd-order base box: its colour rows hold 257 points, more than
    a whole-row workgroup's 256 threads (2 segments per row), and with 64
    threads per colour-pass workgroup its sweeps run 4 segments per row.
    The residual on every point and two sweeps against the oracle's
    multi-colour order (1e-12 relative), the segmented sweeps against the
    unsegmented ones bit for bit."""
    from oracle import oracle as O
    from tests.test_gpu_parity import rel_err, seed_state

    before = oa.sweep_threads(oa.PREC_FP64, 2)
    print("case: long rows", " ".join(LONG_ROWS), flush=True)
    h = oa.Hierarchy(oa.problem_from_argv(LONG_ROWS))
    s = O.Solver.from_argv(LONG_ROWS, sweep_order=O.COLOUR)
    try:
        top = s.levels - 1
        ib = s.level_boxes(top)[0]
        assert h.boxes()[ib]["n"] == [513, 513, 513]
        seed_state(h, s, 6)
        ng = h.residual(top)
        no = s.residual(ib)
        assert abs(ng - no) <= 1e-12 * no, (ng, no)
        R = 1
        g = h.get_field(ib, oa.F_R)[R:-R, R:-R, R:-R]
        o = s.field(ib, O.F_R)[R:-R, R:-R, R:-R]
        assert rel_err(g, o) < 1e-12
        u0 = h.get_field(ib, oa.F_U)
        oa.set_sweep_threads(oa.PREC_FP64, 2, 256)
        h.smooth(top, 2)
        u256 = h.get_field(ib, oa.F_U)
        h.set_field(ib, oa.F_U, u0)
        oa.set_sweep_threads(oa.PREC_FP64, 2, 64)
        ng = h.smooth(top, 2)
        no = s.smooth(ib, 2)
        assert abs(ng - no) <= 1e-11 * no, (ng, no)
        u64 = h.get_field(ib, oa.F_U)
        assert np.array_equal(u64, u256)
        assert rel_err(u64, s.field(ib, O.F_U)) < 1e-12
    finally:
        oa.set_sweep_threads(oa.PREC_FP64, 2, before)
        h.close()
        s.close()
