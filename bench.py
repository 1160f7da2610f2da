"""Benchmark of the GPU FAS puncture solver on one MI355X.

Headline (`value`): solves/s of the plain FP64 path on the largest
single-GPU workload, three punctures with a scalar-field shell
(Punctures_scalar_field; phi = A exp(-(r-r0)^2/sigma^2), A = 0.01, r0 = 10,
sigma = 2; bare masses 0.4, 0.3, 0.3 at positions seeded 2026 within
|x| < 6; L = 256, N = 161, 10 refinement levels, 6th-order FD, nu2 = 20,
tol 1e-10; 139 M points, 10.5 GiB of fields).  A step resets the resident hierarchy to
u = 0, solves it under the reference's stopping rule and computes the ADM
masses (allocation-free entry point).  Geometry, transfer tables and the
field slab are built before the timed region; no host buffers cross PCIe
in a step.

Beside it, on the same line:
  roofline      the FP64 6th-order smoother colour pass (k_smooth<6,1>):
                algorithmic bytes per launch over its HIP-event duration in
                one instrumented solve, against MI355X's 8 TB/s; measured
                HBM traffic from the committed rocprofv3 counter summary
  mixed         the same workload in mixed precision (FP32 correction
                smoothing): solves/s and max |u - u_FP64| over every
                determined point
  binary        the equal-mass binary (N=65, 9 levels, 4th order): one
                request alone in FP64, a batch of 16 in FP64 and in mixed
                precision (engine, one launch per stage for all 16)
  stream        the first requests of the seeded served stream (NumBH
                1..4, N in {33,49,65,97}, 5..9 levels, orders 2/4/6, L 256,
                tol 1e-10): the plain FP64 path (each request solved alone,
                one after the other) against the served engine in FP64 and
                in mixed precision with every request present at t = 0
                (continuous batching); goodput counts only requests that
                reached tol and, in mixed precision, lie within 1e-8 of
                their FP64 solve on every determined point; p50/p99
                latency over those
  cpu_baseline  the reference solver compiled from its sources
                (oracle/_ref/Olliptic.x) on one host core, run in the
                background during the GPU legs on the same workload at
                N = 25 (same punctures, shell, levels, order, nu2, tol),
                timed whole and scaled to N = 161 by the points ratio
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

from ollipticid_amd.workloads import WORKLOADS  # noqa: E402

METRIC = "solves/s at ||du||_inf<=1e-8 vs FP64, p50/p99 latency; smoother HBM GB/s"
HBM_PEAK_GBS = 8000.0
HEADLINE = WORKLOADS["scalar_shell"]
BINARY = WORKLOADS["binary"]
TRAFFIC_PROFILE = "pmc_k_smooth_fp64_scalar_shell.md"
DESCRIPTION = ("three punctures + scalar-field shell (Punctures_scalar_field, A=0.01, r0=10, "
               "sigma=2), masses 0.4/0.3/0.3 at positions seeded 2026 within |x|<6, L=256, "
               "N=161, 10 levels, 6th order, nu2=20, tol 1e-10, plain FP64, u0=0, the "
               "reference's stopping rule; one resident request per step")
CPU_SAMPLE_N = 25


def with_n(argv, n):
    a = list(argv)
    a[a.index("-N") + 1] = str(n)
    return a


def start_cpu_baseline():
    """The reference binary on the headline workload at N = 25, started in
    the background (one core); returns (process, scratch dir, start time)."""
    ref = os.path.join(REPO, "oracle", "_ref", "Olliptic.x")
    if not os.path.exists(ref):
        return None
    tmp = tempfile.mkdtemp()
    argv = with_n(HEADLINE, CPU_SAMPLE_N)
    p = subprocess.Popen([ref] + argv + ["-o", "cpu", "-p1d", "no", "-pint", "no"], cwd=tmp,
                         stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    return p, tmp, time.perf_counter()


def finish_cpu_baseline(job, oa, full_points):
    if job is None:
        return None
    p, tmp, t0 = job
    p.wait()
    t = time.perf_counter() - t0
    try:
        cycles = sum(1 for _ in open(os.path.join(tmp, "cpu", "Loo.gp")))
    except OSError:
        cycles = None
    shutil.rmtree(tmp, ignore_errors=True)
    if p.returncode != 0:
        return None
    boxes, _ = oa.problem_boxes(oa.problem_from_argv(with_n(HEADLINE, CPU_SAMPLE_N)))
    pts = int(sum(np.prod(b["n"]) for b in boxes))
    scale = full_points / pts
    levels = len({b["level"] for b in boxes})
    return {"value": 1.0 / (t * scale), "unit": "solves/s", "cores": 1, "kind": "reference",
            "sample": "the reference solver (oracle/_ref, compiled from its sources) on the "
                      "headline workload at N=%d (%d points, %d V-cycles to its stopping "
                      "rule, setup and ADM included) took %.1f s on 1 host core; scaled by "
                      "the points ratio %.1f to N=161 (%d points)" %
                      (CPU_SAMPLE_N, pts, cycles // levels if cycles else -1, t, scale,
                       full_points)}


def determined_masks(h):
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


def max_diff(h, ref_fields, masks):
    return max(float(np.abs(h.get_field(ib, 0) - ref_fields[ib])[masks[ib]].max())
               for ib in range(h.nbox))


def timed_solves(h, warmup, steps):
    """reset + solve + ADM (no allocation) per step; returns wall, latencies,
    (cycles, residual, adm)"""
    out = None
    for _ in range(warmup):
        h.reset()
        h.solve()
        h.adm_mass_noalloc()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        h.reset()
        cycles, _, res = h.solve()
        adm, _ = h.adm_mass_noalloc()
        lat.append(time.perf_counter() - ts)
        out = (cycles, res, adm)
    return time.perf_counter() - t0, lat, out


def smoother_roofline(oa, prob, order, traffic_profile):
    """One instrumented solve through the engine: HIP events stamped by the
    runtime at the start and end of every smoother launch (on the engine's
    stream); the FP64 colour passes of the given order."""
    e = oa.Engine(graphs=False)
    try:
        assert e.admit(prob, 0)
        e.set_kernel_timing(True)
        while e.step():
            pass
        b, ms, kind = e.kernel_launches()
        e.pop_done()
    finally:
        e.close()
    sel = kind == 2 * order
    n = int(sel.sum())
    avg_ms = float(ms[sel].mean())
    per_launch = float(b[sel].mean())
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    traffic = None
    path = os.path.join(REPO, "profiles", traffic_profile)
    if os.path.isfile(path):
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                traffic = float(line.split(":")[1])
    return {"bound": "hbm", "kernel": "k_smooth<%d,%d> (FP64 colour pass)" % (order, prob.eq),
            "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "traffic_source": "profiles/" + traffic_profile, "launches": n,
            "avg_launch_us": 1e3 * avg_ms, "bytes_per_launch": per_launch,
            "smoother_share_of_launch_time": float(ms[sel].sum() / ms.sum())}


def binary_legs(oa, graphs, warmup, steps):
    """The equal-mass binary: alone in FP64, batches of 16 in FP64 and mixed."""
    prob = oa.problem_from_argv(BINARY)
    h = oa.Hierarchy(prob)
    h.set_graphs(graphs)
    wall, lat, (cycles, res, adm) = timed_solves(h, warmup, steps)
    u64 = [h.get_field(ib, oa.F_U) for ib in range(h.nbox)]
    masks = determined_masks(h)
    out = {"workload": "equal-mass binary N=65, 9 levels, 4th order, nu2=20, tol 1e-11",
           "single_fp64": {"solves_per_s": steps / wall, "p50_ms": 1e3 * float(np.median(lat)),
                           "cycles": cycles, "residual": res}}
    h.close()
    for name, prec in (("batch16_fp64", oa.PREC_FP64), ("batch16_mixed", oa.PREC_MIXED)):
        p = oa.problem_from_argv(BINARY)
        p.precision = prec
        e = oa.Engine(graphs=graphs)
        hs = [oa.Hierarchy(p) for _ in range(16)]

        def batch():
            for t, hh in enumerate(hs):
                assert e.admit_hier(hh, t)
            while e.step():
                pass
            done = sorted(e.pop_hier({t: p for t in range(16)}), key=lambda d: d[0])
            assert len(done) == 16 and all(not isinstance(d[3], int) for d in done)
            return done
        for _ in range(warmup):
            hs = [d[3] for d in batch()]
        t0 = time.perf_counter()
        for _ in range(steps):
            done = batch()
            hs = [d[3] for d in done]
        w = time.perf_counter() - t0
        du = max_diff(hs[0], u64, masks)
        out[name] = {"solves_per_s": 16 * steps / w, "ms_per_batch": 1e3 * w / steps,
                     "cycles": done[0][1], "du_inf_vs_fp64": du}
        for hh in hs:
            hh.close()
        e.close()
    out["batch16_mixed_vs_single_fp64"] = (out["batch16_mixed"]["solves_per_s"] /
                                           out["single_fp64"]["solves_per_s"])
    return out


def stream_legs(oa, nreq, graphs):
    """The first nreq requests of the seeded served stream (seed 2026)."""
    from ollipticid_amd import serve

    argvs, p64 = serve.make_stream(nreq, 2026, precision="fp64")
    _, pmx = serve.make_stream(nreq, 2026, precision="mixed")
    keep = set(range(nreq))
    # the plain FP64 path: each request alone, one after another
    t0 = time.perf_counter()
    single = []
    for p in p64:
        try:
            with oa.Hierarchy(p) as h:
                cycles, _, res = h.solve()
                single.append(res <= 1e-10)
        except oa.OllError:
            single.append(False)
    wall_single = time.perf_counter() - t0
    out = {"requests": nreq, "seed": 2026,
           "plain_fp64_alone": {"wall_s": wall_single, "solves_per_s": nreq / wall_single,
                                "goodput_per_s": sum(single) / wall_single,
                                "reached_tol": int(sum(single))}}
    ref = {}
    for name, probs in (("served_fp64", p64), ("served_mixed", pmx)):
        res, wall, extra = serve.serve(probs, [0.0] * nreq, 160 << 30, graphs=graphs, keep=keep)
        good, lat = 0, []
        refused = sum(1 for r in res if r.get("refused"))
        worst = 0.0
        for i, r in enumerate(res):
            if r is None or "hier" not in r:
                continue
            ok = r["residual"] <= 1e-10
            if name == "served_fp64":
                if ok:
                    ref[i] = ([r["hier"].get_field(ib, oa.F_U) for ib in range(r["hier"].nbox)],
                              determined_masks(r["hier"]))
            elif ok and i in ref:
                d = max_diff(r["hier"], *ref[i])
                worst = max(worst, d)
                ok = d <= 1e-8
            else:
                ok = False
            r["hier"].close()
            if ok:
                good += 1
                lat.append(r["done"] - r["arrival"])
        out[name] = {"wall_s": wall, "solves_per_s": nreq / wall, "goodput_per_s": good / wall,
                     "good": good, "refused": refused,
                     "p50_ms": 1e3 * float(np.percentile(lat, 50)) if lat else None,
                     "p99_ms": 1e3 * float(np.percentile(lat, 99)) if lat else None,
                     "mean_active": extra["mean_active"], "max_active": extra["max_active"]}
        if name == "served_mixed":
            out[name]["du_inf_vs_fp64_max"] = worst
    out["served_mixed_goodput_vs_plain_fp64"] = (out["served_mixed"]["goodput_per_s"] /
                                                out["plain_fp64_alone"]["goodput_per_s"])
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--stream", type=int, default=48,
                    help="served-stream requests in the stream leg (0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-legs", action="store_true",
                    help="headline and roofline only (no mixed, binary or stream legs)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="issue every launch from the host instead of replaying captured "
                         "V-cycle graphs (profilers that cannot trace graph launches)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; the served stream shards across "
              "GPUs with no exchange, so no multi-GPU figure is produced", file=sys.stderr)
        sys.exit(2)

    import ollipticid_amd as oa

    graphs = not args.no_graphs
    job = None if args.no_cpu_baseline else start_cpu_baseline()

    prob = oa.problem_from_argv(HEADLINE)
    h = oa.Hierarchy(prob)
    h.set_graphs(graphs)
    wall, lat, (cycles, res, adm) = timed_solves(h, args.warmup, args.steps)
    info = {"boxes": h.nbox, "levels": h.levels,
            "points": int(sum(np.prod(b["n"]) for b in h.boxes())),
            "hbm_bytes": int(h.device_bytes())}
    u64 = [h.get_field(ib, oa.F_U) for ib in range(h.nbox)]
    masks = determined_masks(h)
    h.close()
    roof = smoother_roofline(oa, prob, 6, TRAFFIC_PROFILE)
    out = {
        "metric": METRIC,
        "value": args.steps / wall,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "graphs": graphs,
        "data": "synthetic",
        "config": {"workload": DESCRIPTION, "batch": 1, **info, "parallelism": "single GPU"},
        "p50_ms": 1e3 * float(np.percentile(lat, 50)),
        "p99_ms": 1e3 * float(np.percentile(lat, 99)),
        "cycles": cycles,
        "residual": res,
        "adm_mass": list(adm),
        "roofline": roof,
    }
    if not args.no_legs:
        pm = oa.problem_from_argv(HEADLINE)
        pm.precision = oa.PREC_MIXED
        hm = oa.Hierarchy(pm)
        hm.set_graphs(graphs)
        wall_m, lat_m, (cm, rm, am) = timed_solves(hm, args.warmup, args.steps)
        du = max_diff(hm, u64, masks)
        hm.close()
        out["mixed"] = {"solves_per_s": args.steps / wall_m, "ms_per_step": 1e3 * wall_m / args.steps,
                        "dtype": "f32 correction smoothing + f64 defect/update/transfers/coarsest",
                        "cycles": cm, "residual": rm, "adm_mass": list(am),
                        "du_inf_vs_fp64": du, "vs_fp64": (args.steps / wall_m) / out["value"]}
        assert du <= 1e-8, du
        out["binary"] = binary_legs(oa, graphs, args.warmup, args.steps)
        if args.stream > 0:
            out["stream"] = stream_legs(oa, args.stream, graphs)
    cpu = finish_cpu_baseline(job, oa, info["points"])
    if cpu:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
