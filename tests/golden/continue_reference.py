"""Continues a reference run past the reference's stopping rule.

The reference stops when the top level's norm stops decreasing
(elliptic.cc:1265-1330); on some hierarchies the base box is then not yet
converged (the equal-mass binary, N=65, 9 levels: its base box moves by
2e-9 over the next hundred V-cycles).  This script replays the reference
run with the oracle in the reference's own lexicographic order, checks that
it reproduces the stored reference run (same cycle count, every stored u
value to 1e-14), continues the same V-cycle for `extra` more cycles,
records how far any stored u value still moved over the last `tail` of
them, and writes the result as a fixture of the same format (u of every
box, subsampled by the stored stride; the ADM masses):

    python tests/golden/continue_reference.py cfg2_allconv cfg2_continued
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle as O  # noqa: E402


def stored_u(s, st):
    return [s.field(ib, O.F_U)[::st, ::st, ::st].copy() for ib in range(s.nbox)]


def main(src, dst, extra=150, tail=20):
    with open(os.path.join(HERE, src + ".json")) as f:
        meta = json.load(f)
    arrays = dict(np.load(os.path.join(HERE, src + ".npz")))
    st = meta["stride"]
    s = O.Solver.from_argv(meta["argv"], sweep_order=O.LEX)
    cycles, hist, res = s.solve()
    assert cycles == meta["cycles"], (cycles, meta["cycles"])
    seen = {}
    worst = 0.0
    for ib, b in enumerate(s.boxes()):
        lv = b["level"]
        i = seen.get(lv, 0)
        seen[lv] = i + 1
        ref = arrays["u_l%d_b%d" % (lv, i)]
        worst = max(worst, float(np.abs(s.field(ib, O.F_U)[::st, ::st, ::st] - ref).max()))
    print("replayed %d cycles; max |u_oracle - u_reference| %.3e" % (cycles, worst), flush=True)
    assert worst <= 1e-14
    top = s.levels - 1
    for c in range(extra):
        if c == extra - tail:
            before = stored_u(s, st)
        r = s.fasn(top)
    moved = max(float(np.abs(a - b).max()) for a, b in zip(stored_u(s, st), before))
    print("continued %d cycles; top-level norm %.3e; largest change over the last %d: %.3e"
          % (extra, r, tail, moved), flush=True)
    out, boxes, seen = {}, [], {}
    punct, surf = s.adm_mass()
    for ib, b in enumerate(s.boxes()):
        lv = b["level"]
        i = seen.get(lv, 0)
        seen[lv] = i + 1
        out["u_l%d_b%d" % (lv, i)] = s.field(ib, O.F_U)[::st, ::st, ::st].copy()
        boxes.append({"level": lv, "box": i, "n": b["n"], "origin": b["origin"], "dx": b["d"],
                      "surface_adm": float(surf[ib])})
    new = {"name": dst, "argv": meta["argv"], "stride": st, "levels": meta["levels"],
           "boxes": boxes, "adm_mass": [float(x) for x in punct],
           "cycles": cycles + extra, "reference_cycles": cycles,
           "replay_max_diff": worst, "moved_over_last_%d_cycles" % tail: moved,
           "note": "the reference run %s (%d V-cycles to its stopping rule) continued %d "
                   "V-cycles by the oracle in lexicographic order, which reproduces the "
                   "reference run to %.1e; u stored lexicographic (k,j,i), full precision"
                   % (src, cycles, extra, worst)}
    np.savez_compressed(os.path.join(HERE, dst + ".npz"), **out)
    with open(os.path.join(HERE, dst + ".json"), "w") as f:
        json.dump(new, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1], sys.argv[2])
