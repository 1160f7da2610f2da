"""Autotuner of the launch-shape table: threads per workgroup of the
smoothers' colour passes, per precision and order.

For each (precision, order) it times the colour passes of the equal-mass
binary workload at that order (N=65, 9 levels, nu2=20; the engine's
per-launch HIP events, launches issued one by one) with every candidate
workgroup size, twice: as a batch of requests and as one request alone
(the launches of a lightly loaded server).  The rate of each is the
algorithmic bytes of the updated points over the passes' device time.  A
size replaces the 256-thread default only when it is faster than the
default in both regimes by at least MIN_GAIN; of those, the one whose worse
regime gains most.  The colour passes' results do not depend on the size:
the tuner only changes speed.  It writes ollipticid_amd/launch_shapes.json, which the
binding applies when the library loads:

    python -m ollipticid_amd.autotune [--batch 16] [--cycles 2]

Prints one JSON line per measurement and the chosen table.
"""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

PUNC = ["-eq", "Punctures", "-B", "Robin", "-BA", "0", "-Bn", "1"]
BINARY = PUNC + ["-NumBH", "2", "-bhmp1", "0.487209", "-bhy1", "5.5",
                 "-bhpx1", "-0.0901099", "-bhpy1", "-0.000703975",
                 "-bhmp2", "0.487209", "-bhy2", "-5.5",
                 "-bhpx2", "0.0901099", "-bhpy2", "0.000703975",
                 "-L", "1000", "-N", "65", "-l", "9", "-nu2", "20", "-s", "1e-11"]
CANDIDATES = (64, 128, 256, 512)
DEFAULT = 256
MIN_GAIN = 1.02


def measure(oa, order, precision, batch, cycles):
    """Algorithmic GB/s of the colour passes over `cycles` V-cycles of
    `batch` requests, and their launch count."""
    p = oa.problem_from_argv(BINARY + ["-O", str(order)])
    p.precision = precision
    p.max_cycles = cycles
    e = oa.Engine(graphs=False)
    try:
        for t in range(batch):
            assert e.admit(p, t)
        e.set_kernel_timing(True)
        while e.step():
            pass
        launches, ms, nbytes = e.kernel_timing()
        e.pop_done()
    finally:
        e.close()
    return nbytes / (ms * 1e-3) / 1e9 if ms > 0 else 0.0, launches


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--cycles", type=int, default=2)
    ap.add_argument("--orders", default="2,4,6,8")
    ap.add_argument("--precisions", default="fp64,mixed")
    ap.add_argument("--out", default=os.path.join(HERE, "launch_shapes.json"))
    args = ap.parse_args()

    from ollipticid_amd import native as oa

    oa.load()
    table, measured = {}, {}
    for pname in args.precisions.split(","):
        prec = oa.PRECISION_NAMES[pname]
        for order in (int(o) for o in args.orders.split(",")):
            rates = {}
            for threads in CANDIDATES:
                oa.set_sweep_threads(prec, order, threads)
                rb, nb = measure(oa, order, prec, args.batch, args.cycles)
                rs, ns = measure(oa, order, prec, 1, args.cycles)
                rates[threads] = (rb, rs)
                print(json.dumps({"precision": pname, "order": order, "threads": threads,
                                  "batch_GBs": round(rb, 1), "alone_GBs": round(rs, 1)}),
                      flush=True)
            base = rates[DEFAULT]
            gain = {t: min(r[0] / base[0], r[1] / base[1]) for t, r in rates.items()}
            best = max(gain, key=gain.get)
            if gain[best] < MIN_GAIN:
                best = DEFAULT
            oa.set_sweep_threads(prec, order, best)
            table.setdefault(pname, {})[str(order)] = best
            measured.setdefault(pname, {})[str(order)] = {
                str(t): {"batch": round(r[0], 1), "alone": round(r[1], 1)} for t, r in rates.items()}
    out = {"sweep_threads": table,
           "measured_colour_pass_GBs": measured,
           "workload": "equal-mass binary N=65 l=9 nu2=20 at each order, batch of %d and alone, "
                       "%d V-cycles" % (args.batch, args.cycles),
           "rule": "a size replaces %d only when faster in both regimes by %.0f%% or more"
                   % (DEFAULT, 100 * (MIN_GAIN - 1))}
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({"sweep_threads": table}), flush=True)


if __name__ == "__main__":
    main()
