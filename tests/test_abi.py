"""CPU checks of the drop-in boundary: the HIP library loads, exports every
function include/olliptic_amd.h declares, the Python binding declares the
same set, and the host-only geometry entry points reproduce the reference's
box tables (tests/golden)."""
import ctypes
import glob
import json
import os
import re

import numpy as np
import pytest

import ollipticid_amd as oa
from ollipticid_amd import native

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "olliptic_amd.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_]+\s*\**\s*(oll_[a-z_0-9]+)\s*\(",
                                 text, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "oll_hier_create" in names and "oll_solve" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(native.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_functions()) == set(native.SIGNATURES)


@pytest.mark.parametrize("path", sorted(p for p in glob.glob(os.path.join(REPO, "tests", "golden", "*.json"))
                                         if not os.path.basename(p).startswith(("out_", "io_"))))
def test_box_tables_match_reference(path):
    meta = json.load(open(path))
    boxes, mgrid = oa.problem_boxes(oa.problem_from_argv(meta["argv"]))
    assert [(b["level"], b["n"]) for b in boxes] == [(b["level"], b["n"]) for b in meta["boxes"]]
    for b, r in zip(boxes, meta["boxes"]):
        # the fixtures print 12 significant digits
        for a in range(3):
            assert abs(b["origin"][a] - r["origin"][a]) <= 1e-11 * max(1.0, abs(r["origin"][a]))
            assert abs(b["d"][a] - r["dx"][a]) <= 1e-11 * r["dx"][a]


def test_geometry_errors_are_codes_not_exits():
    p = oa.problem_from_argv(["-eq", "Punctures", "-B", "Robin", "-N", "17", "-L", "10"])
    with pytest.raises(oa.OllError) as e:
        oa.problem_boxes(p)
    assert e.value.code == -2 and "enough grid-points" in str(e.value)


def test_cli_parsing_rules():
    # repeated flag keeps the last value; unknown flags ignored; -bh* positional
    p = oa.problem_from_argv(["-eq", "Punctures", "-N", "33", "-N", "65", "-foo", "1",
                              "-NumBH", "2", "-bhmp2", "0.7", "-bhy1", "3.5"])
    assert p.N == 65 and p.num_bh == 2
    assert p.bh[1].m == 0.7 and p.bh[0].y == 3.5 and p.bh[0].m == 0.0
    # fewer than two arguments: nothing parsed (interface.cc:373)
    q = oa.problem_from_argv(["-N"])
    assert q.N == 30


def test_equation_names_and_robin_values():
    """Equation names as the reference's -eq takes them; the Robin value is
    the reference's boundary_coeff (operator.cc:665, 744, 791) unless -BA
    sets it; the library accepts every equation and rejects others."""
    base = ["-B", "Robin", "-NumBH", "1", "-bhmp1", "0.5", "-L", "20", "-N", "25", "-l", "1"]
    want = {"Punctures": (oa.EQ_PUNCTURES, 0.0), "Punctures_scalar_field": (oa.EQ_PUNCTURES_SCALAR, 0.0),
            "Poisson": (oa.EQ_POISSON, 0.0), "Brill-waves": (oa.EQ_BRILL, 1.0),
            "Punctures_EM": (oa.EQ_PUNCTURES_EM, 1.0)}
    for name, (eq, a) in want.items():
        p = oa.problem_from_argv(["-eq", name] + base + ["-O", "2"])
        assert (p.eq, p.robin_a) == (eq, a), name
        assert oa.problem_footprint(p) > 0
    assert oa.problem_from_argv(["-eq", "Brill-waves", "-BA", "0.5"] + base).robin_a == 0.5
    with pytest.raises(ValueError):
        oa.problem_from_argv(["-eq", "Bowen-York"] + base)
    p = oa.problem_from_argv(["-eq", "Poisson"] + base)
    p.eq = 5
    with pytest.raises(oa.OllError):
        oa.problem_footprint(p)
    p.eq = oa.EQ_POISSON
    p.initial_guess = oa.GUESS_APPROX
    with pytest.raises(oa.OllError):
        oa.problem_footprint(p)
    # Punctures_EM: charges parsed; FP64 only; six more fields
    p = oa.problem_from_argv(["-eq", "Punctures_EM", "-bhqp1", "0.25", "-O", "2"] + base)
    assert p.bh[0].q == 0.25
    em_bytes = oa.problem_footprint(p)
    p.eq = oa.EQ_BRILL
    assert em_bytes > oa.problem_footprint(p)
    p.eq = oa.EQ_PUNCTURES_EM
    p.order = 4
    assert oa.problem_footprint(p) > 0
    p.precision = oa.PREC_MIXED
    with pytest.raises(oa.OllError):
        oa.problem_footprint(p)


def test_sweep_threads_table():
    """The launch-shape table: multiples of 64 from 64 to 512 per precision
    and order; anything else refused (no GPU needed)."""
    before = oa.sweep_threads(oa.PREC_MIXED, 6)
    try:
        oa.set_sweep_threads(oa.PREC_MIXED, 6, 128)
        assert oa.sweep_threads(oa.PREC_MIXED, 6) == 128
        assert oa.sweep_threads(oa.PREC_FP64, 6) in (64, 128, 256, 512)
        for prec, order, threads in ((oa.PREC_FP64, 4, 96), (oa.PREC_FP64, 4, 1024),
                                     (oa.PREC_FP64, 3, 256), (7, 4, 256)):
            with pytest.raises(oa.OllError):
                oa.set_sweep_threads(prec, order, threads)
    finally:
        oa.set_sweep_threads(oa.PREC_MIXED, 6, before)


def test_colour_order_matches_oracle():
    """The host geometry's per-box colour order (oll_problem_boxes) is the
    oracle's: a 2nd-order case with a box swept colour 1 first, and 24
    requests of the served stream (2nd, 4th and 6th order; two-colour boxes
    on and off their parent's lattice)."""
    from oracle import oracle as O
    from ollipticid_amd import serve
    from tests.test_gpu_parity import O2_FLIP

    rng = np.random.default_rng(2026)
    argvs = [O2_FLIP] + [serve.request_argv(rng) for _ in range(24)]
    flipped = 0
    for argv in argvs:
        try:
            boxes, _ = oa.problem_boxes(oa.problem_from_argv(argv))
        except oa.OllError:  # a box list the reference cannot set up either
            continue
        got = [b["colour_order"] for b in boxes]
        s = O.Solver.from_argv(argv, sweep_order=O.COLOUR)
        assert got == [s.colour_flip(i) for i in range(s.nbox)], argv
        flipped += sum(got)
    assert 1 in [b["colour_order"] for b in oa.problem_boxes(oa.problem_from_argv(O2_FLIP))[0]]
    assert flipped > 1
