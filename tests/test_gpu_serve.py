"""The serving driver end to end on a short seeded stream of served-workload
requests with Poisson arrivals: requests overlap on the device, and every
result is the request solved alone in the same precision, bit for bit, and
within 1e-8 of its FP64 solve on every determined point where both reached
the tolerance."""
import numpy as np
import pytest

import ollipticid_amd as oa
from ollipticid_amd import serve

pytestmark = pytest.mark.gpu


def test_served_stream_equals_alone():
    argvs, probs = serve.make_stream(96, 11)
    small = [i for i, p in enumerate(probs) if oa.problem_footprint(p) < 150e6][:10]
    assert len(small) >= 6
    argvs = [argvs[i] for i in small]
    probs = [probs[i] for i in small]
    rng = np.random.default_rng(3)
    arrivals = list(np.cumsum(rng.exponential(1.0 / 200.0, len(probs))) - 0.005)
    out, wall, extra = serve.serve(probs, arrivals, 8 << 30, keep=set(range(len(probs))))
    summary = serve.summarize(out, wall, extra)
    print("served:", summary, flush=True)
    assert extra["max_active"] > 1
    assert all("error" not in r for r in out)
    worst, n = serve.check_against_alone(argvs, probs, out, list(range(len(probs))))
    print("max |u - u_FP64| over determined points of", n, "converged requests:", worst, flush=True)
    assert n >= 1
    assert worst <= 1e-8
