"""The batched engine: requests of different N, level count, order, equation
and precision solved together give, request by request, exactly what each
gives alone (every u value of every box, bit for bit), also when they arrive
while others are mid-solve; admission respects the HBM budget."""
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
    # mixed precision member
    PUNC + ["-NumBH", "2", "-bhmp1", "0.5", "-bhx1", "1.5", "-bhpy1", "0.1", "-bhmp2", "0.5",
            "-bhx2", "-1.5", "-bhpy2", "-0.1", "-L", "40", "-N", "25", "-l", "3", "-O", "4",
            "-nu2", "3", "-s", "1e-10", "-prec", "mixed"],
]


def fields(h):
    return [h.get_field(ib, oa.F_U) for ib in range(h.nbox)]


def alone(argv):
    with oa.Hierarchy(oa.problem_from_argv(argv)) as h:
        cycles, hist, res = h.solve()
        adm, _ = h.adm_mass()
        return cycles, res, adm, fields(h)


def check_same(tag, ref, got):
    c, r, a, u = ref
    gc, gr, h = got
    assert gc == c, (tag, gc, c)
    assert gr == r, (tag, gr, r)
    np.testing.assert_array_equal(h.adm_mass()[0], a)
    for ib, (x, y) in enumerate(zip(fields(h), u)):
        np.testing.assert_array_equal(x, y, err_msg="request %d box %d" % (tag, ib))


def run_to_end(e, problems, got, arrivals=None):
    """Steps the engine; arrivals maps step index -> list of tags admitted
    just before that step."""
    step = 0
    while True:
        for tag in (arrivals or {}).get(step, []):
            assert e.admit(problems[tag], tag)
        active = e.step()
        for tag, cycles, res, h in e.pop_hier(problems):
            got[tag] = (cycles, res, h)
        step += 1
        pending = [s for s in (arrivals or {}) if s >= step]
        if active == 0 and not pending:
            return
        assert step < 400


def test_mix_equals_alone():
    ref = [alone(a) for a in MIX]
    print("alone:", [(r[0], r[1]) for r in ref], flush=True)
    problems = {t: oa.problem_from_argv(a) for t, a in enumerate(MIX)}
    e = oa.Engine()
    for tag in problems:
        assert e.admit(problems[tag], tag)
    got = {}
    run_to_end(e, problems, got)
    assert sorted(got) == list(range(len(MIX)))
    for tag in got:
        check_same(tag, ref[tag], got[tag])


def chunk_argv(t, mixed):
    argv = PUNC + ["-NumBH", "1", "-bhmp1", "0.5", "-bhx1", "%.2f" % (0.3 + 0.05 * t), "-L", "64",
                   "-N", "65", "-l", "4", "-O", "4", "-nu2", "2", "-s", "1e-10"]
    return argv + (["-prec", "mixed"] if mixed else [])


def test_chunked_blocks_equal_alone():
    """Launches over many requests give each transfer and whole-row block
    several chunks of points or rows (16 requests with 67^3 boxes: 18.8k
    prolongation blocks of one chunk, joined 4 to a block; the 8 FP64 and 8
    mixed members' defect and residual blocks 2 to a block); every request
    still gives bit for bit its solve alone."""
    argvs = [chunk_argv(t, t >= 8) for t in range(16)]
    ref = [alone(a) for a in argvs]
    problems = {t: oa.problem_from_argv(a) for t, a in enumerate(argvs)}
    e = oa.Engine()
    for tag in problems:
        assert e.admit(problems[tag], tag)
    got = {}
    run_to_end(e, problems, got)
    assert sorted(got) == list(range(len(argvs)))
    for tag in got:
        check_same(tag, ref[tag], got[tag])


def test_staggered_arrivals_equal_alone():
    """Continuous batching: requests join while others are mid-solve."""
    ref = [alone(a) for a in MIX]
    problems = {t: oa.problem_from_argv(a) for t, a in enumerate(MIX)}
    e = oa.Engine()
    got = {}
    run_to_end(e, problems, got, arrivals={0: [0, 3], 2: [1], 3: [4], 5: [2]})
    assert sorted(got) == list(range(len(MIX)))
    for tag in got:
        check_same(tag, ref[tag], got[tag])


def test_resident_readmission():
    """A solved hierarchy moved back into the engine solves again from u = 0
    to the same result, without being rebuilt."""
    ref = alone(MIX[1])
    p = oa.problem_from_argv(MIX[1])
    e = oa.Engine()
    h = oa.Hierarchy(p)
    for rnd in range(2):
        assert e.admit_hier(h, 7)
        got = {}
        run_to_end(e, {7: p}, got)
        check_same(7, ref, got[7])
        h = got[7][2]
    h.close()


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


def test_kernel_launch_log_matches_totals():
    """The per-launch smoother log holds one entry per timed launch and adds
    up to the engine's totals; FP64 and FP32 colour passes are told apart."""
    e = oa.Engine(graphs=False)
    try:
        for t, argv in enumerate(MIX[:2] + MIX[4:]):
            assert e.admit(oa.problem_from_argv(argv), t)
        e.step()
        e.set_kernel_timing(True)
        e.step()
        n, ms, nbytes = e.kernel_timing()
        b, lms, kind = e.kernel_launches()
        assert n == len(b) == len(lms) == len(kind) > 0
        assert abs(b.sum() - nbytes) <= 1e-9 * nbytes
        assert abs(lms.sum() - ms) <= 1e-6 * ms + 1e-6
        # kind = 2 * order + precision (0 FP64, 1 FP32 correction)
        assert set((kind % 2).tolist()) == {0, 1}
        assert set((kind // 2).tolist()) <= {2, 4, 6, 8}
        assert (b > 0).all() and (lms > 0).all()
    finally:
        e.close()
