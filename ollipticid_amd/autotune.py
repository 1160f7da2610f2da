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

For FP64 it also times the fused double-buffered sweeps (two colour passes
per launch on tiles of ty x tz interior rows) for each order with every
candidate tile shape and workgroup size, in the same two regimes, against
the chosen colour-by-colour size: the best shape per order goes into the
table, and the fused sweeps are switched on only when they beat the colour
passes in both regimes by MIN_GAIN at every order measured.

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
FUSED_CANDIDATES = ((4, 4, 256), (4, 8, 256), (8, 8, 256), (8, 8, 512), (16, 16, 512))
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
    fused_shape, fused_measured, fused_gain = {}, {}, {}
    if "fp64" in args.precisions.split(","):
        prec = oa.PRECISION_NAMES["fp64"]
        for order in (int(o) for o in args.orders.split(",")):
            base = measured["fp64"][str(order)][str(table["fp64"][str(order)])]
            rates = {}
            oa.set_fused_sweeps(True)
            try:
                for shape in FUSED_CANDIDATES:
                    oa.set_fused_shape(order, *shape)
                    rb, _ = measure(oa, order, prec, args.batch, args.cycles)
                    rs, _ = measure(oa, order, prec, 1, args.cycles)
                    rates[shape] = (rb, rs)
                    print(json.dumps({"precision": "fp64", "order": order, "fused_shape": shape,
                                      "batch_GBs": round(rb, 1), "alone_GBs": round(rs, 1)}),
                          flush=True)
            finally:
                oa.set_fused_sweeps(False)
            gain = {t: min(r[0] / base["batch"], r[1] / base["alone"]) for t, r in rates.items()}
            best = max(gain, key=gain.get)
            oa.set_fused_shape(order, *best)
            fused_shape[str(order)] = list(best)
            fused_gain[str(order)] = round(gain[best], 3)
            fused_measured[str(order)] = {"%dx%d/%d" % t: {"batch": round(r[0], 1),
                                                             "alone": round(r[1], 1)}
                                          for t, r in rates.items()}
    fused_on = bool(fused_gain) and min(fused_gain.values()) >= MIN_GAIN
    oa.set_fused_sweeps(fused_on)
    out = {"sweep_threads": table,
           "measured_colour_pass_GBs": measured,
           "fused_sweeps": fused_on,
           "fused_shape": fused_shape,
           "fused_gain_vs_colour_passes": fused_gain,
           "measured_fused_GBs": fused_measured,
           "workload": "equal-mass binary N=65 l=9 nu2=20 at each order, batch of %d and alone, "
                       "%d V-cycles" % (args.batch, args.cycles),
           "rule": "a size replaces %d only when faster in both regimes by %.0f%% or more"
                   % (DEFAULT, 100 * (MIN_GAIN - 1))}
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({"sweep_threads": table, "fused_sweeps": fused_on, "fused_shape": fused_shape}),
          flush=True)


if __name__ == "__main__":
    main()
