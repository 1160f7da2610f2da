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
