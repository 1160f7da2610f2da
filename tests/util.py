"""Shared helpers of the test-suite."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")


def load_golden(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        meta = json.load(f)
    arrays = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
    return meta, arrays


def golden_u(meta, arrays, boxes):
    """Golden u per box (same order as `boxes`), subsampled by the stored stride."""
    out = []
    seen = {}
    for b in boxes:
        lv = b["level"]
        i = seen.get(lv, 0)
        seen[lv] = i + 1
        out.append(arrays["u_l%d_b%d" % (lv, i)])
    return out


def determined_mask(box, mgrid):
    """Points whose value the discrete problem fixes.  The Robin shell of the
    coarsened multigrid levels (levels < mgrid) only carries FAS correction
    states: their absolute values depend on the iteration history, not on
    the solution, so field comparisons leave them out."""
    n = box["n"]
    m = np.ones((n[2], n[1], n[0]), dtype=bool)
    if box["level"] < mgrid:
        R = box["R"]
        m[:] = False
        m[R:-R, R:-R, R:-R] = True
    return m
