"""The serving driver end to end on requests drawn from the served stream
itself (seed 2026, the benchmark's stream: NumBH 1..4, N in {33, 49, 65,
97}, 5..9 levels, orders 2/4/6), with Poisson arrivals: a seeded sample of
ten plus three whose multi-colour V-cycle used to diverge (merged boxes off
the coarser lattice) and one whose box merging the reference cannot set up.
Requests overlap on the device; every served result is the request solved
alone in the same precision, bit for bit; every request but the refused one
reaches tol 1e-10 and lies within 1e-8 of its FP64 solve on every
determined point; the refused one is answered with OLL_EINVAL."""
import numpy as np
import pytest

import ollipticid_amd as oa
from ollipticid_amd import serve

pytestmark = pytest.mark.gpu

FORMERLY_DIVERGENT = [66, 102, 448]
REFUSED = 19


def test_served_stream_equals_alone():
    argvs, probs = serve.make_stream(512, 2026)
    sample = sorted(int(i) for i in np.random.default_rng(5).choice(512, 10, replace=False))
    picks = [i for i in sample if i != REFUSED] + FORMERLY_DIVERGENT + [REFUSED]
    print("requests:", picks, flush=True)
    argvs = [argvs[i] for i in picks]
    probs = [probs[i] for i in picks]
    rng = np.random.default_rng(3)
    arrivals = list(np.cumsum(rng.exponential(1.0 / 50.0, len(probs))) - 0.02)
    out, wall, extra = serve.serve(probs, arrivals, 64 << 30, keep=set(range(len(probs))))
    summary = serve.summarize(out, wall, extra)
    print("served:", {k: v for k, v in summary.items() if k != "engine"}, flush=True)
    assert extra["max_active"] > 1
    assert out[-1].get("refused") and out[-1]["error"] == -1
    served = list(range(len(probs) - 1))
    assert all("error" not in out[i] for i in served)
    assert all(out[i]["residual"] <= 1e-10 for i in served), [out[i]["residual"] for i in served]
    worst, n = serve.check_against_alone(argvs, probs, out, served)
    print("max |u - u_FP64| over determined points of", n, "requests:", worst, flush=True)
    assert n == len(served)
    assert worst <= 1e-8
