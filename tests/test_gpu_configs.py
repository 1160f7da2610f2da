"""GPU parity on the shapes of every benchmark configuration.

* the equal-mass binary (N=65, 9 levels, 4th order), the same binary at
  8th order and the three punctures with the scalar-field shell at 6th
  order (the N=129 and N=161 workloads reduced to N=25), against reference
  runs converged on every level (tests/golden/cfg3r_o8, cfg5r_o6; for the
  binary, whose base box the reference's stopping rule leaves 2e-9 short
  of convergence, its run continued past that rule by the oracle in the
  reference's order, tests/golden/cfg2_continued): every determined point,
  the puncture ADM masses and the surface ADM masses, in FP64 and in mixed
  precision, solved with the workloads' own sweep counts (the discrete
  solution does not depend on them);
* the two full-size workloads (N=129, 31 M points; N=161, 139 M points),
  where no reference run exists: FP64 and mixed precision solved to the
  workload's tolerance, and the mixed solution within 1e-8 of the FP64 one
  on every determined point.

Bars: 1e-9 absolute on u against the reference, the printed digits (9) of
the ADM masses, 1e-8 between mixed and FP64.
"""
import numpy as np
import pytest

import ollipticid_amd as oa
from ollipticid_amd.workloads import WORKLOADS
from tests.util import determined_mask, golden_u, load_golden

pytestmark = pytest.mark.gpu

ADM_RTOL = 5.01e-9  # half a unit of the 9th printed digit


def with_sweeps(argv, nu1=3, nu2=20):
    """The workloads' own sweep counts (the reference's default nu1 = 3,
    nu2 = 20) in place of the fixture run's."""
    a = list(argv)
    for flag, v in (("-nu1", nu1), ("-nu2", nu2)):
        if flag in a:
            a[a.index(flag) + 1] = str(v)
        else:
            a += [flag, str(v)]
    return a


def solve(argv, precision, stop=None, tol=None, max_cycles=200, patience=10):
    p = oa.problem_from_argv(argv)
    p.precision = precision
    if stop is not None:
        p.stop_rule = stop
        p.max_cycles = max_cycles
        p.patience = patience
    if tol is not None:
        p.tol = tol
    h = oa.Hierarchy(p)
    cycles, hist, res = h.solve()
    return h, cycles, res


def compare_to_reference(h, meta, arrays, label):
    st = meta["stride"]
    worst = 0.0
    per_level = {}
    for ib, (b, ref) in enumerate(zip(h.boxes(), golden_u(meta, arrays, h.boxes()))):
        m = determined_mask(b, h.mgrid)[::st, ::st, ::st]
        u = h.get_field(ib, oa.F_U)[::st, ::st, ::st]
        assert u.shape == ref.shape, (label, ib, u.shape, ref.shape)
        d = float(np.abs(u - ref)[m].max())
        per_level[b["level"]] = max(per_level.get(b["level"], 0.0), d)
        worst = max(worst, d)
    adm, surf = h.adm_mass()
    print("%s: max |u - u_ref| %.3e, ADM %s; per level %s" %
          (label, worst, adm, " ".join("%d:%.1e" % kv for kv in sorted(per_level.items()))),
          flush=True)
    assert worst <= 1e-9, (label, worst)
    np.testing.assert_allclose(adm, meta["adm_mass"], rtol=ADM_RTOL)
    # the surface integral of a coarsened multigrid level (below the base
    # box) samples its Robin shell, which carries FAS correction states
    # rather than values the discrete problem fixes (they depend on the
    # sweep order): those boxes' masses are compared with the oracle in
    # test_gpu_parity only
    ref_surf = [b["surface_adm"] for b in meta["boxes"]]
    assert len(ref_surf) == len(surf)
    for ib, (b, g, r) in enumerate(zip(h.boxes(), surf, ref_surf)):
        if b["level"] >= h.mgrid:
            assert abs(g - r) <= ADM_RTOL * abs(r), (label, "surface ADM", ib, g, r)


@pytest.mark.parametrize("name", ["cfg2_continued", "cfg3r_o8", "cfg5r_o6"])
@pytest.mark.parametrize("prec", [oa.PREC_FP64, oa.PREC_MIXED], ids=["fp64", "mixed"])
def test_workload_shapes_match_reference(name, prec):
    """The binary (N=65, 4th order), the 8th-order binary and the 6th-order
    scalar-field shell at N=25, every level converged by the reference;
    solved here with the workloads' 3 + 20 sweeps to 1e-13
    (OLL_STOP_TOLERANCE)."""
    meta, arrays = load_golden(name)
    argv = with_sweeps(meta["argv"])
    print("case:", name, " ".join(argv), flush=True)
    h, cycles, res = solve(argv, prec, stop=oa.STOP_TOLERANCE, tol=1e-13)
    print("cycles", cycles, "res", res, flush=True)
    assert res <= 1e-12
    compare_to_reference(h, meta, arrays, "%s prec %d" % (name, prec))


def determined_diff(h1, h2):
    worst = 0.0
    for ib, b in enumerate(h1.boxes()):
        m = determined_mask(b, h1.mgrid)
        d = np.abs(h1.get_field(ib, oa.F_U) - h2.get_field(ib, oa.F_U))[m]
        worst = max(worst, float(d.max()))
    return worst


@pytest.mark.parametrize("name", ["binary_hires", "scalar_shell"])
def test_full_size_workload_mixed_matches_fp64(name):
    """Full-size N=129 8th-order binary and N=161 scalar-field shell: both
    precisions reach the workload's tolerance; mixed within 1e-8 of FP64."""
    argv = WORKLOADS[name]
    tol = float(argv[argv.index("-s") + 1])
    print("case:", name, " ".join(argv), flush=True)
    h64, c64, r64 = solve(argv, oa.PREC_FP64, stop=oa.STOP_TOLERANCE, max_cycles=120)
    print("fp64 cycles", c64, "res", r64, flush=True)
    hm, cm, rm = solve(argv, oa.PREC_MIXED, stop=oa.STOP_TOLERANCE, max_cycles=120)
    print("mixed cycles", cm, "res", rm, flush=True)
    assert r64 <= tol and rm <= tol
    d = determined_diff(hm, h64)
    print("%s: max |u_mixed - u_FP64| %.3e" % (name, d), flush=True)
    assert d <= 1e-8
    np.testing.assert_allclose(hm.adm_mass()[0], h64.adm_mass()[0], rtol=1e-9)
    hm.close()
    h64.close()
