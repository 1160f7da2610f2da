"""oll_adm_mass_noalloc (the ADM masses without device allocation, staging
in the library's per-GPU scratch array) against oll_adm_mass on the same
solved hierarchies: bit for bit, on two calls in a row, the second on a
smaller hierarchy (fewer boxes and punctures), and the first hierarchy
again after it (the scratch keeps the last call's contents)."""
import numpy as np
import pytest

import ollipticid_amd as oa

pytestmark = pytest.mark.gpu

PUNC = ["-eq", "Punctures", "-B", "Robin", "-BA", "0", "-Bn", "1"]
LARGE = PUNC + ["-NumBH", "3", "-bhmp1", "0.4", "-bhx1", "2.5", "-bhpy1", "0.1",
                "-bhmp2", "0.35", "-bhx2", "-2.0", "-bhy2", "1.5", "-bhsz2", "0.05",
                "-bhmp3", "0.3", "-bhy3", "-2.5", "-bhpx3", "-0.1",
                "-L", "64", "-N", "33", "-l", "4", "-O", "4", "-nu2", "6", "-s", "1e-10"]
SMALL = PUNC + ["-NumBH", "1", "-bhmp1", "1.0", "-bhpx1", "0.2", "-L", "64", "-N", "25",
                "-l", "2", "-O", "2", "-nu2", "6", "-s", "1e-10"]


def solved(argv):
    h = oa.Hierarchy(oa.problem_from_argv(argv))
    h.solve()
    return h


def test_noalloc_adm_equals_allocating_path():
    print("case: large", " ".join(LARGE), flush=True)
    big = solved(LARGE)
    print("case: small", " ".join(SMALL), flush=True)
    small = solved(SMALL)
    assert small.nbox < big.nbox and small.problem.num_bh < big.problem.num_bh
    for h in (big, small, big):
        p0, s0 = h.adm_mass()
        p1, s1 = h.adm_mass_noalloc()
        assert np.array_equal(p0, p1), (p0, p1)
        assert np.array_equal(s0, s1), (s0, s1)
        assert all(np.isfinite(p1)) and all(np.isfinite(s1))
    big.close()
    small.close()
