"""Fused FP64 sweeps (k_smooth_fused: two colour passes per launch on tiles
of interior rows, u double-buffered between the u and r fields) against the
same solve swept colour by colour: the cycle count, the norm history, the
ADM masses and every u value of every box agree bit for bit, for every
stencil order, the four equations the fused sweeps serve (Punctures_EM is
swept colour by colour either way), even and odd sweep counts (an odd count
ends with one colour-by-colour sweep), two-colour boxes swept in either
colour order, several tile shapes, and the batched engine."""
import numpy as np
import pytest

import ollipticid_amd as oa
from ollipticid_amd import native
from tests.util import load_golden

pytestmark = pytest.mark.gpu

CASES = ["small_o2_4bh", "small_o4_offgrid", "small_o6_merge", "small_o8", "small_sf_3bh",
         "small_brill_o8", "small_poisson_o2", "small_em_o4"]


def set_sweeps(argv, nu1, nu2):
    a = list(argv)
    for flag, v in (("-nu1", nu1), ("-nu2", nu2)):
        if flag in a:
            a[a.index(flag) + 1] = str(v)
        else:
            a += [flag, str(v)]
    return a


def run(argv, fused, max_cycles=6):
    native.set_fused_sweeps(fused)
    try:
        p = oa.problem_from_argv(argv)
        p.stop_rule = oa.STOP_TOLERANCE
        p.tol = 1e-300  # a fixed number of cycles
        p.max_cycles = max_cycles
        with oa.Hierarchy(p) as h:
            cycles, hist, res = h.solve()
            return cycles, np.asarray(hist), res, h.adm_mass()[0], [h.get_field(ib, oa.F_U)
                                                                for ib in range(h.nbox)]
    finally:
        native.set_fused_sweeps(False)


def check_equal(tag, a, b):
    assert a[0] == b[0], (tag, "cycles", a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1], err_msg=tag + " history")
    assert a[2] == b[2], (tag, "residual", a[2], b[2])
    np.testing.assert_array_equal(a[3], b[3], err_msg=tag + " ADM")
    assert len(a[4]) == len(b[4])
    for ib, (x, y) in enumerate(zip(a[4], b[4])):
        np.testing.assert_array_equal(x, y, err_msg="%s box %d" % (tag, ib))


@pytest.mark.parametrize("sweeps", [(2, 4), (3, 5)], ids=["even", "odd"])
@pytest.mark.parametrize("name", CASES)
def test_fused_equals_colour_by_colour(name, sweeps):
    meta, _ = load_golden(name)
    argv = set_sweeps(meta["argv"], *sweeps)
    print("case:", name, "nu1 %d nu2 %d" % sweeps, flush=True)
    ref = run(argv, False)
    got = run(argv, True)
    print("cycles", got[0], "res", got[2], flush=True)
    check_equal(name, ref, got)


@pytest.mark.parametrize("shape", [(1, 1, 64), (3, 5, 128), (16, 16, 512), (8, 2, 256)])
def test_fused_tile_shapes(shape):
    """The tile shape and workgroup size change nothing."""
    meta, _ = load_golden("small_o6_merge")
    argv = set_sweeps(meta["argv"], 2, 4)
    ref = run(argv, False, max_cycles=3)
    saved = native.fused_shape(6)
    print("shape", shape, "default", saved, flush=True)
    native.set_fused_shape(6, *shape)
    try:
        got = run(argv, True, max_cycles=3)
    finally:
        native.set_fused_shape(6, *saved)
    check_equal("shape %s" % (shape,), ref, got)


def test_fused_shape_rejects_bad_values():
    with pytest.raises(native.OllError):
        native.set_fused_shape(5, 8, 8, 256)
    with pytest.raises(native.OllError):
        native.set_fused_shape(4, 0, 8, 256)
    with pytest.raises(native.OllError):
        native.set_fused_shape(4, 8, 8, 100)


def test_fused_engine_batch():
    """Requests of several orders and equations batched in the engine."""
    names = ["small_o2_4bh", "small_o4_offgrid", "small_sf_3bh", "small_o8"]
    argvs = [set_sweeps(load_golden(n)[0]["argv"], 3, 4) for n in names]

    def batch(fused):
        native.set_fused_sweeps(fused)
        try:
            problems = {}
            for t, a in enumerate(argvs):
                p = oa.problem_from_argv(a)
                p.stop_rule = oa.STOP_TOLERANCE
                p.tol = 1e-300
                p.max_cycles = 4
                problems[t] = p
            e = oa.Engine()
            for t, p in problems.items():
                assert e.admit(p, t)
            out = {}
            for _ in range(100):
                active = e.step()
                for tag, cycles, res, h in e.pop_hier(problems):
                    out[tag] = (cycles, res, [h.get_field(ib, oa.F_U) for ib in range(h.nbox)])
                    h.close()
                if active == 0:
                    break
            e.close()
            return out
        finally:
            native.set_fused_sweeps(False)

    ref = batch(False)
    got = batch(True)
    assert sorted(ref) == sorted(got) == list(range(len(argvs)))
    for t in ref:
        assert ref[t][0] == got[t][0] and ref[t][1] == got[t][1], (names[t], ref[t][:2], got[t][:2])
        for ib, (x, y) in enumerate(zip(ref[t][2], got[t][2])):
            np.testing.assert_array_equal(x, y, err_msg="%s box %d" % (names[t], ib))
