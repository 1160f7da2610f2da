"""The served-stream generator: seeded, within the stated ranges, punctures
at least 2 apart, every request a valid hierarchy (host-only checks)."""
import numpy as np

import ollipticid_amd as oa
from ollipticid_amd import serve


def test_stream_is_seeded_and_in_range():
    a1, p1 = serve.make_stream(64, 5)
    a2, _ = serve.make_stream(64, 5)
    a3, _ = serve.make_stream(64, 6)
    assert a1 == a2 and a1 != a3
    for p in p1:
        assert 1 <= p.num_bh <= 4
        assert p.N in serve.N_CHOICES and p.order in serve.ORDER_CHOICES
        assert 5 <= p.levels <= 9 and p.tol == 1e-10 and p.L == serve.L_DOMAIN
        assert p.precision == oa.PREC_MIXED
        xs = [np.array([p.bh[i].x, p.bh[i].y, p.bh[i].z]) for i in range(p.num_bh)]
        for i in range(p.num_bh):
            b = p.bh[i]
            assert 0.3 <= b.m <= 0.6
            assert all(-8 <= v <= 8 for v in (b.x, b.y, b.z))
            assert all(-0.2 <= v <= 0.2 for v in (b.px, b.py, b.pz, b.sx, b.sy, b.sz))
            for j in range(i):
                assert np.linalg.norm(xs[i] - xs[j]) >= 2.0


def test_stream_requests_build():
    _, probs = serve.make_stream(32, 7)
    for p in probs:
        assert oa.problem_footprint(p) > 0
