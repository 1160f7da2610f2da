"""Served stream: puncture requests of uneven N, level count and order arrive
over time and are solved by one engine with continuous batching.

Requests join the running batch at the next V-cycle boundary as soon as
their HBM footprint fits the budget; every engine step runs one V-cycle of
every active request (one launch per stage for all of them); finished
requests leave.  Latency is arrival -> result.

The synthetic stream follows the served-stream workload: NumBH uniform in
1..4, bare masses U(0.3, 0.6), positions U(-8, 8)^3 with pairwise separation
>= 2, momentum and spin components U(-0.2, 0.2), N in {33, 49, 65, 97},
refinement levels 5..9, order in {2, 4, 6}, tol 1e-10, Robin boundary,
L = 256, seeded; Poisson arrivals.

    python -m ollipticid_amd.serve --requests 512 --seed 2026 --rate auto
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if os.path.dirname(HERE) not in sys.path:
    sys.path.insert(0, os.path.dirname(HERE))

L_DOMAIN = 256.0
N_CHOICES = (33, 49, 65, 97)
ORDER_CHOICES = (2, 4, 6)
# the reference's stall stop (res < previous res) ends most of these solves
# around 1e-5; served solves run to tol with these limits instead
MAX_CYCLES = 60
PATIENCE = 8


def request_argv(rng):
    """One request as a reference command line."""
    nbh = int(rng.integers(1, 5))
    pos = []
    while len(pos) < nbh:
        x = rng.uniform(-8.0, 8.0, 3)
        if all(np.linalg.norm(x - y) >= 2.0 for y in pos):
            pos.append(x)
    f = lambda v: "%.17g" % float(v)  # noqa: E731  (exact round trip through strtod)
    argv = ["-eq", "Punctures", "-B", "Robin", "-BA", "0", "-Bn", "1", "-NumBH", str(nbh)]
    for i, x in enumerate(pos, start=1):
        m = rng.uniform(0.3, 0.6)
        p = rng.uniform(-0.2, 0.2, 3)
        sp = rng.uniform(-0.2, 0.2, 3)
        argv += ["-bhmp%d" % i, f(m), "-bhx%d" % i, f(x[0]), "-bhy%d" % i, f(x[1]),
                 "-bhz%d" % i, f(x[2]), "-bhpx%d" % i, f(p[0]), "-bhpy%d" % i, f(p[1]),
                 "-bhpz%d" % i, f(p[2]), "-bhsx%d" % i, f(sp[0]), "-bhsy%d" % i, f(sp[1]),
                 "-bhsz%d" % i, f(sp[2])]
    argv += ["-L", f(L_DOMAIN), "-N", str(int(rng.choice(N_CHOICES))),
             "-l", str(int(rng.integers(5, 10))), "-O", str(int(rng.choice(ORDER_CHOICES))),
             "-s", "1e-10"]
    return argv


def stream_problem(argv, precision="mixed", stop="tolerance", guess="zero"):
    import ollipticid_amd as oa

    p = oa.problem_from_argv(argv)
    p.precision = oa.PREC_MIXED if precision == "mixed" else oa.PREC_FP64
    p.initial_guess = oa.GUESS_APPROX if guess == "approx" else oa.GUESS_ZERO
    if stop == "tolerance":
        p.stop_rule = oa.STOP_TOLERANCE
        p.max_cycles = MAX_CYCLES
        p.patience = PATIENCE
    return p


def make_stream(n, seed, precision="mixed", stop="tolerance", guess="zero"):
    """n request command lines (seeded) and their problems."""
    rng = np.random.default_rng(seed)
    argvs = [request_argv(rng) for _ in range(n)]
    return argvs, [stream_problem(a, precision, stop, guess) for a in argvs]


def serve(probs, arrivals, hbm_budget, graphs=True, max_active=0, log=None, keep=()):
    """Runs the stream; arrivals[i] is request i's arrival time in seconds
    from the start.  Returns per-request dicts (arrival, admitted, done,
    cycles, residual, adm; the solved Hierarchy for requests in `keep`) and
    the wall time."""
    import ollipticid_amd as oa

    order = sorted(range(len(probs)), key=lambda i: arrivals[i])
    e = oa.Engine(hbm_budget=hbm_budget, graphs=graphs)
    out = [None] * len(probs)
    waiting = []
    nxt = 0
    active = 0
    steps = 0
    batch_sizes = []
    t0 = time.perf_counter()
    while nxt < len(order) or waiting or active:
        now = time.perf_counter() - t0
        while nxt < len(order) and arrivals[order[nxt]] <= now:
            waiting.append(order[nxt])
            nxt += 1
        while waiting and (max_active <= 0 or active < max_active):
            i = waiting[0]
            try:
                if not e.admit(probs[i], i):
                    break
            except oa.OllError as err:
                # a problem the solver cannot set up (the reference's grid
                # set-up fails on it too): answered at once with the error
                waiting.pop(0)
                now = time.perf_counter() - t0
                out[i] = {"arrival": arrivals[i], "admitted": now, "done": now, "cycles": 0,
                          "residual": float("inf"), "error": err.code, "refused": True}
                continue
            waiting.pop(0)
            active += 1
            out[i] = {"arrival": arrivals[i], "admitted": time.perf_counter() - t0}
        if active == 0:
            if nxt < len(order):
                time.sleep(max(0.0, arrivals[order[nxt]] - (time.perf_counter() - t0)))
            continue
        batch_sizes.append(active)
        active = e.step()
        steps += 1
        for tag, cycles, res, h in e.pop_hier({i: probs[i] for i in range(len(probs))}):
            r = out[tag]
            r["cycles"] = cycles
            r["residual"] = res
            if isinstance(h, int):
                r["error"] = h
            else:
                r["adm"] = list(h.adm_mass()[0])
                if tag in keep:
                    r["hier"] = h
                else:
                    h.close()
            r["done"] = time.perf_counter() - t0
        if log and steps % 50 == 0:
            print("step %d t=%.1fs active %d waiting %d arrived %d/%d" %
                  (steps, time.perf_counter() - t0, active, len(waiting), nxt, len(order)),
                  file=log, flush=True)
    wall = time.perf_counter() - t0
    prof = e.profile()
    e.close()
    return out, wall, {"steps": steps, "mean_active": float(np.mean(batch_sizes)) if batch_sizes else 0.0,
                       "max_active": int(max(batch_sizes)) if batch_sizes else 0, "engine": prof}


def summarize(out, wall, extra, tol=1e-10):
    """goodput: requests that reached tol per second of wall time, with the
    latency percentiles over those; the completed figures count every
    request answered (reached tol, stopped above it, failed or refused)"""
    good = [r for r in out if "error" not in r and r["residual"] <= tol]
    lat_good = np.array([r["done"] - r["arrival"] for r in good])
    lat_all = np.array([r["done"] - r["arrival"] for r in out])
    pct = lambda a, q: 1e3 * float(np.percentile(a, q)) if len(a) else None  # noqa: E731
    return {"requests": len(out), "wall_s": wall,
            "goodput_per_s": len(good) / wall, "reached_tol": len(good),
            "p50_ms": pct(lat_good, 50), "p99_ms": pct(lat_good, 99),
            "completed_per_s": len(out) / wall,
            "completed_p50_ms": pct(lat_all, 50), "completed_p99_ms": pct(lat_all, 99),
            "above_tol": sum(1 for r in out if "error" not in r and r["residual"] > tol),
            "failed": sum(1 for r in out if "error" in r and not r.get("refused")),
            "refused": sum(1 for r in out if r.get("refused")),
            "mean_cycles": float(np.mean([r["cycles"] for r in out if not r.get("refused")])),
            **extra}


def u_fields(h):
    import ollipticid_amd as oa

    return [h.get_field(ib, oa.F_U) for ib in range(h.nbox)]


def determined(h):
    """Per box: the points the discrete problem fixes (all but the Robin
    shell of the coarsened multigrid levels, which carries FAS states)."""
    out = []
    for b in h.boxes():
        n = b["n"]
        m = np.ones((n[2], n[1], n[0]), dtype=bool)
        if b["level"] < h.mgrid:
            R = b["R"]
            m[:] = False
            m[R:-R, R:-R, R:-R] = True
        out.append(m)
    return out


def check_against_alone(argvs, probs, out, picks, tol=1e-10):
    """For the picked (kept) requests: the served result is the request
    solved alone in the same precision, bit for bit (cycles, residual, every
    u value).  Returns max |u - u_FP64| over every determined point of the
    picked requests where both the served and the FP64 solve reached tol,
    and how many those were."""
    import ollipticid_amd as oa

    worst, n = 0.0, 0
    for i in picks:
        r = out[i]
        if "hier" not in r:  # the request failed (NaN divergence): nothing to compare
            continue
        served = u_fields(r["hier"])
        masks = determined(r["hier"])
        with oa.Hierarchy(probs[i]) as h:
            cycles, _, res = h.solve()
            alone = u_fields(h)
        assert r["cycles"] == cycles and r["residual"] == res, (i, r["cycles"], cycles, r["residual"], res)
        assert all(np.array_equal(a, b) for a, b in zip(served, alone)), i
        with oa.Hierarchy(stream_problem(argvs[i], "fp64")) as h:
            _, _, res64 = h.solve()
            ref = u_fields(h)
        if res <= tol and res64 <= tol:
            n += 1
            worst = max(worst, max(float(np.abs(a - b)[m].max()) for a, b, m in zip(served, ref, masks)))
        r["hier"].close()
        del r["hier"]
    return worst, n


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--requests", type=int, default=512)
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--rate", default="auto",
                    help="Poisson arrival rate in requests/s, or 'auto': 0.8 x the capacity "
                         "measured with every request present at t = 0")
    ap.add_argument("--precision", choices=("mixed", "fp64"), default="mixed")
    ap.add_argument("--stop", choices=("tolerance", "reference"), default="tolerance",
                    help="tolerance: V-cycles until tol (at most %d, or %d without a new best); "
                         "reference: the reference's stall rule" % (MAX_CYCLES, PATIENCE))
    ap.add_argument("--guess", choices=("zero", "approx"), default="zero",
                    help="initial u: the reference's 0, or the analytic puncture approximation")
    ap.add_argument("--hbm-gib", type=float, default=160.0, help="admission budget")
    ap.add_argument("--check", type=int, default=4, help="requests re-solved alone to check")
    args = ap.parse_args()

    argvs, probs = make_stream(args.requests, args.seed, args.precision, args.stop, args.guess)
    budget = int(args.hbm_gib * (1 << 30))
    capacity = None
    if args.rate == "auto":
        out, wall, extra = serve(probs, [0.0] * len(probs), budget, log=sys.stderr)
        capacity = len(probs) / wall
        rate = 0.8 * capacity
    else:
        rate = float(args.rate)
    rng = np.random.default_rng(args.seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(probs)))
    arrivals -= arrivals[0]
    picks = [int(i) for i in np.random.default_rng(args.seed + 2).choice(len(probs), args.check,
                                                                          replace=False)]
    out, wall, extra = serve(probs, list(arrivals), budget, log=sys.stderr, keep=set(picks))
    res = summarize(out, wall, extra)
    res.update({"seed": args.seed, "rate_per_s": rate, "capacity_per_s": capacity,
                "precision": args.precision, "guess": args.guess, "hbm_budget_gib": args.hbm_gib,
                "L": L_DOMAIN,
                "stop": args.stop, "max_cycles": MAX_CYCLES, "patience": PATIENCE,
                "errors": sum(1 for r in out if "error" in r)})
    if picks:
        res["checked_alone"] = picks
        res["du_inf_vs_fp64_converged"], res["checked_converged"] = check_against_alone(
            argvs, probs, out, picks)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
