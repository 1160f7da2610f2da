"""The batched engine: requests of different N, level count, order and
equation solved together give, request by request, exactly what each gives
alone; admission respects the HBM budget."""
import numpy as np
import pytest

import ollipticid_amd as oa

pytestmark = pytest.mark.gpu

PUNC = ["-eq", "Punctures", "-B", "Robin"]
MIX = [
    PUNC + ["-NumBH", "1", "-bhmp1", "1.0", "-bhpx1", "0.2", "-L", "256", "-N", "33", "-l", "4",
            "-O", "2", "-nu2", "3", "-s", "1e-10"],
    PUNC + ["-NumBH", "2", "-bhmp1", "0.5", "-bhx1", "1.5", "-bhpy1", "0.1", "-bhmp2", "0.5",
            "-bhx2", "-1.5", "-bhpy2", "-0.1", "-L", "40", "-N", "25", "-l", "3", "-O", "4",
            "-nu2", "3", "-s", "1e-10"],
    ["-eq", "Punctures_scalar_field", "-B", "Robin", "-psc", "2", "-phi0", "0.02", "-a2", "1",
     "-r0", "10", "-sigma", "4", "-NumBH", "1", "-bhmp1", "0.5", "-bhx1", "0.7", "-bhpy1", "0.1",
     "-L", "64", "-N", "49", "-l", "2", "-O", "6", "-nu2", "3", "-s", "1e-10"],
    PUNC + ["-NumBH", "3", "-bhmp1", "0.3", "-bhx1", "3", "-bhmp2", "0.4", "-bhy2", "-2.5",
            "-bhmp3", "0.35", "-bhz3", "2.2", "-bhpx3", "0.1", "-L", "96", "-N", "49", "-l", "5",
            "-O", "4", "-nu2", "3", "-s", "1e-10"],
]


def alone(argv):
    with oa.Hierarchy(oa.problem_from_argv(argv)) as h:
        cycles, hist, res = h.solve()
        adm, _ = h.adm_mass()
    return cycles, res, adm


def test_mix_equals_alone():
    ref = [alone(a) for a in MIX]
    print("alone:", ref, flush=True)
    e = oa.Engine()
    for tag, a in enumerate(MIX):
        assert e.admit(oa.problem_from_argv(a), tag)
    got = {}
    for _ in range(200):
        if e.step() == 0:
            break
    for tag, cycles, res, adm in e.pop_done():
        got[tag] = (cycles, res, adm)
    assert sorted(got) == list(range(len(MIX)))
    for tag, (c, r, a) in enumerate(ref):
        gc, gr, ga = got[tag]
        assert gc == c, (tag, gc, c)
        assert gr == r, (tag, gr, r)
        np.testing.assert_array_equal(ga[:len(a)], a)


def test_budget_admission():
    p = oa.problem_from_argv(MIX[0])
    need = oa.problem_footprint(p)
    e = oa.Engine(hbm_budget=int(need * 2.5))
    assert e.admit(p, 1) and e.admit(p, 2)
    assert not e.admit(p, 3)            # would exceed the budget
    while e.step():
        pass
    assert len(e.pop_done()) == 2
    assert e.admit(p, 3)                # memory freed
    while e.step():
        pass
    (tag, cycles, res, adm), = e.pop_done()
    assert tag == 3 and res <= 1e-10
