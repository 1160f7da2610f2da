"""Benchmark: solves/s of the plain FP64 GPU FAS solve on the equal-mass
binary workload (NumBH=2, bare masses 0.487209 at y=+-5.5, momenta
+-(-0.0901099,-0.000703975,0), Robin, L=1000, N=65, 9 refinement levels,
4th-order FD, nu2=20, tol 1e-11), one MI355X.

A step is one solve: reset u to the initial guess 0, V-cycles until the
reference's stopping rule, ADM masses.  The hierarchy (geometry, transfer
tables, coefficients) is built before the timed region and stays resident.

Prints one JSON line (see the contract in the task) with the dominant
kernel's roofline (the multi-colour smoother, HIP events around each launch
of one extra instrumented solve) and the reference CPU solver timed on the
host on a bounded sample.
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

METRIC = "solves/s at ||du||_inf<=1e-8 vs FP64, p50/p99 latency; smoother HBM GB/s"
PUNC = ["-eq", "Punctures", "-B", "Robin", "-BA", "0", "-Bn", "1"]
BINARY = PUNC + ["-NumBH", "2", "-bhmp1", "0.487209", "-bhy1", "5.5",
                 "-bhpx1", "-0.0901099", "-bhpy1", "-0.000703975",
                 "-bhmp2", "0.487209", "-bhy2", "-5.5",
                 "-bhpx2", "0.0901099", "-bhpy2", "0.000703975",
                 "-L", "1000", "-N", "65", "-l", "9", "-O", "4", "-nu2", "20", "-s", "1e-11"]
HBM_PEAK_GBS = 8000.0


def cpu_baseline(argv, full_cycles):
    """Reference solver (compiled from its sources into oracle/_ref) on the
    host: first V-cycle only (tolerance above the first residual), scaled to
    the full solve's cycle count.  Falls back to the oracle port."""
    ref = os.path.join(REPO, "oracle", "_ref", "Olliptic.x")
    sample_argv = list(argv)
    sample_argv[sample_argv.index("-s") + 1] = "1e-1"
    if os.path.exists(ref):
        tmp = tempfile.mkdtemp()
        try:
            t0 = time.perf_counter()
            subprocess.run([ref] + sample_argv + ["-o", "cpu", "-p1d", "no", "-pint", "no",
                                                  "-so", "no"], cwd=tmp, check=True,
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
            t = time.perf_counter() - t0
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        kind = "reference"
    else:
        from oracle import oracle as O
        s = O.Solver.from_argv(sample_argv, max_cycles=1)
        t0 = time.perf_counter()
        s.solve()
        s.adm_mass()
        t = time.perf_counter() - t0
        kind = "port"
    per_solve = t * full_cycles
    return {"value": 1.0 / per_solve, "unit": "solves/s", "cores": 1, "kind": kind,
            "sample": "first V-cycle of the same workload (setup + 1 of %d cycles + ADM) "
                      "took %.2f s on 1 host core; scaled x%d" % (full_cycles, t, full_cycles)}


def batched(prob, B, cycles_alone, adm_alone, graphs):
    """B copies of the workload resident together, V-cycles in lockstep (one
    launch per stage covers all B).  Each copy must reproduce the single solve
    exactly."""
    import ollipticid_amd as oa

    e = oa.Engine(graphs=graphs)
    for t in range(B):
        assert e.admit(prob, t)
    t0 = time.perf_counter()
    while e.step():
        pass
    done = e.pop_done()
    wall = time.perf_counter() - t0
    assert len(done) == B
    for tag, cycles, res, adm in done:
        assert cycles == cycles_alone and adm[:2] == list(adm_alone), (tag, cycles, adm)
    e.close()
    return {"batch": B, "solves_per_s": B / wall, "ms_per_batch": 1e3 * wall,
            "identical_to_single": True}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-graphs", action="store_true",
                    help="issue every launch from the host instead of replaying a captured "
                         "V-cycle graph (profilers that cannot trace graph launches)")
    ap.add_argument("--batch", type=int, default=8,
                    help="also solve this many copies together in one engine batch")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; the served stream shards across "
              "GPUs with no exchange (see DESIGN.md), so no multi-GPU figure is produced",
              file=sys.stderr)
        sys.exit(2)

    import ollipticid_amd as oa

    prob = oa.problem_from_argv(BINARY)
    h = oa.Hierarchy(prob)
    h.set_graphs(not args.no_graphs)

    def step():
        h.reset()
        cycles, hist, res = h.solve()
        adm, _ = h.adm_mass()
        return cycles, res, adm

    for _ in range(args.warmup):
        step()
    lat = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ts = time.perf_counter()
        cycles, res, adm = step()
        lat.append(time.perf_counter() - ts)
    wall = time.perf_counter() - t0

    # one instrumented solve: HIP events around every smoother launch
    h.set_kernel_timing(True)
    step()
    launches, kms, kbytes = h.kernel_timing()
    h.set_kernel_timing(False)
    avg_ms = kms / launches
    achieved = (kbytes / launches) / (avg_ms * 1e-3) / 1e9

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
        "graphs": not args.no_graphs,
        "data": "synthetic",
        "config": {"workload": "equal-mass binary: NumBH=2, m=0.487209 at y=+-5.5, "
                               "P=+-(-0.0901099,-0.000703975,0), Robin, L=1000, N=65, "
                               "l=9, O4, nu2=20, tol 1e-11, plain FP64",
                   "boxes": h.nbox, "levels": h.levels,
                   "points": int(sum(np.prod(b["n"]) for b in h.boxes())),
                   "parallelism": "single GPU"},
        "p50_ms": 1e3 * float(np.percentile(lat, 50)),
        "p99_ms": 1e3 * float(np.percentile(lat, 99)),
        "cycles": cycles,
        "coarse_sweeps_per_solve": h.last_solve_coarse_sweeps(),
        "residual": res,
        "adm_mass": adm,
        "du_inf_vs_fp64": 0.0,
        "roofline": {"bound": "hbm", "kernel": "k_smooth<4,0> (multi-colour GS-Newton pass)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "launches": launches, "avg_launch_us": 1e3 * avg_ms,
                     "bytes_per_launch": kbytes / launches},
    }
    if args.batch > 1:
        out["batched"] = batched(prob, args.batch, cycles, adm, not args.no_graphs)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(BINARY, cycles)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
