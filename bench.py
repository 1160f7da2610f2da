"""Benchmark: solves/s of the GPU FAS puncture solver on the equal-mass
binary workload (NumBH=2, bare masses 0.487209 at y=+-5.5, momenta
+-(-0.0901099,-0.000703975,0), Robin, L=1000, N=65, 9 refinement levels,
4th-order FD, nu2=20, tol 1e-11), one MI355X.

A step is one batch: B copies of the request, each with its own hierarchy
resident in HBM (geometry, transfer tables and fields built before the timed
region), are reset to the initial guess u = 0, solved together by the engine
(V-cycles in lockstep, one launch per stage covering all B) until each meets
the reference's stopping rule, and their ADM masses are computed.  `value`
is solves/s of that plain FP64 path.  Beside it:
  single    the same request solved alone (one hierarchy, graph-replayed
            V-cycles): the plain FP64 GPU path without batching;
  mixed     the batched path with FP32 correction smoothing
            (OLL_PREC_MIXED), its max |u - u_FP64| over every determined
            point of every box, and the same roofline entry for its FP32
            smoother;
  mixed_approx  the mixed batched path started from the analytic puncture
            approximation instead of u = 0 (OLL_GUESS_APPROX): fewer
            V-cycles to the same solution, with its max |u - u_FP64|;
  roofline  the dominant kernel (the FP64 multi-colour smoother pass): HIP
            events around each of its launches in one extra instrumented
            batch;
  cpu_baseline  the reference CPU solver, compiled from its sources into
            oracle/_ref, timed on the host on a bounded sample.
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
SINGLE_STEPS = 5


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


class Batch:
    """B resident hierarchies of one problem, solved together per step."""

    def __init__(self, oa, prob, B, graphs):
        self.oa = oa
        self.engine = oa.Engine(graphs=graphs)
        self.hiers = [oa.Hierarchy(prob) for _ in range(B)]
        self.problems = {t: prob for t in range(B)}
        self.B = B

    def step(self):
        for t, h in enumerate(self.hiers):
            assert self.engine.admit_hier(h, t)
        while self.engine.step():
            pass
        done = sorted(self.engine.pop_hier(self.problems), key=lambda d: d[0])
        assert len(done) == self.B and all(not isinstance(d[3], int) for d in done), done
        self.hiers = [d[3] for d in done]
        adms = [h.adm_mass()[0] for h in self.hiers]
        return [(d[1], d[2], adm) for d, adm in zip(done, adms)]

    def close(self):
        for h in self.hiers:
            h.close()
        self.engine.close()


def timed_batches(batch, warmup, steps):
    for _ in range(warmup):
        batch.step()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ts = time.perf_counter()
        results = batch.step()
        lat.append(time.perf_counter() - ts)
    wall = time.perf_counter() - t0
    return wall, lat, results


def measured_traffic(name):
    """HBM bytes per launch of a smoother from its committed hardware-counter
    summary (rocprofv3 FETCH_SIZE and WRITE_SIZE of the same workload and
    batch, scaled by a device-to-device copy of known size); None if absent."""
    path = os.path.join(REPO, "profiles", name)
    if not name or not os.path.isfile(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":")[1])
    return None


def roofline(batch, kernel, traffic_profile):
    """One more batch with HIP events around every smoother launch."""
    batch.engine.set_kernel_timing(True)
    batch.step()
    launches, kms, kbytes = batch.engine.kernel_timing()
    batch.engine.set_kernel_timing(False)
    avg_ms = kms / launches
    achieved = (kbytes / launches) / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(traffic_profile), "traffic_source": "profiles/" + traffic_profile,
            "launches": launches, "avg_launch_us": 1e3 * avg_ms,
            "bytes_per_launch": kbytes / launches}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=16,
                    help="requests solved together per step")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-graphs", action="store_true",
                    help="issue every launch from the host instead of replaying captured "
                         "V-cycle graphs (profilers that cannot trace graph launches)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only; the served stream shards across "
              "GPUs with no exchange (see DESIGN.md), so no multi-GPU figure is produced",
              file=sys.stderr)
        sys.exit(2)

    import ollipticid_amd as oa

    graphs = not args.no_graphs
    prob = oa.problem_from_argv(BINARY)

    # the plain FP64 path, one request at a time
    h = oa.Hierarchy(prob)
    h.set_graphs(graphs)

    def single():
        h.reset()
        cycles, hist, res = h.solve()
        adm, _ = h.adm_mass()
        return cycles, res, adm

    single()
    lat1 = []
    for _ in range(SINGLE_STEPS):
        ts = time.perf_counter()
        cycles, res, adm = single()
        lat1.append(time.perf_counter() - ts)
    u64 = [h.get_field(ib, oa.F_U) for ib in range(h.nbox)]
    masks = determined_masks(h)
    single_out = {"solves_per_s": 1.0 / float(np.mean(lat1)), "ms_per_solve": 1e3 * float(np.mean(lat1)),
                  "p50_ms": 1e3 * float(np.percentile(lat1, 50)),
                  "p99_ms": 1e3 * float(np.percentile(lat1, 99)), "steps": SINGLE_STEPS,
                  "coarse_sweeps_per_solve": h.last_solve_coarse_sweeps()}
    info = {"boxes": h.nbox, "levels": h.levels,
            "points": int(sum(np.prod(b["n"]) for b in h.boxes()))}
    h.close()

    # the step: a batch of B resident FP64 requests
    batch = Batch(oa, prob, args.batch, graphs)
    wall, lat, results = timed_batches(batch, args.warmup, args.steps)
    for c, r, a in results:
        assert c == cycles and r == res and list(a) == list(adm), (c, r, a)
    roof = roofline(batch, "k_smooth<4,0> (FP64 multi-colour GS-Newton pass)",
                    "pmc_k_smooth_fp64_binary_batch16.md" if args.batch == 16 else "")
    batch.close()

    value = args.batch * args.steps / wall
    out = {
        "metric": METRIC,
        "value": value,
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
        "config": {"workload": "equal-mass binary: NumBH=2, m=0.487209 at y=+-5.5, "
                               "P=+-(-0.0901099,-0.000703975,0), Robin, L=1000, N=65, "
                               "l=9, O4, nu2=20, tol 1e-11, plain FP64; a step solves "
                               "%d copies, each with its own resident hierarchy" % args.batch,
                   "batch": args.batch, **info, "parallelism": "single GPU"},
        "p50_ms": 1e3 * float(np.percentile(lat, 50)),
        "p99_ms": 1e3 * float(np.percentile(lat, 99)),
        "cycles": cycles,
        "residual": res,
        "adm_mass": list(adm),
        "du_inf_vs_fp64": 0.0,
        "identical_to_single": True,
        "roofline": roof,
        "single": single_out,
    }

    if not args.no_mixed:
        pm = oa.problem_from_argv(BINARY)
        pm.precision = oa.PREC_MIXED
        mb = Batch(oa, pm, args.batch, graphs)
        wall_m, lat_m, res_m = timed_batches(mb, args.warmup, args.steps)
        du = 0.0
        for ib in range(len(u64)):
            d = np.abs(mb.hiers[0].get_field(ib, oa.F_U) - u64[ib])[masks[ib]]
            du = max(du, float(d.max()))
        roof_m = roofline(mb, "k_smooth32<4> (FP32 correction colour pass)",
                          "pmc_k_smooth32_mixed_binary_batch16.md" if args.batch == 16 else "")
        mb.close()
        mixed_value = args.batch * args.steps / wall_m
        out["mixed"] = {
            "solves_per_s": mixed_value, "dtype": "f32 smoothing + f64 defect/update/transfers",
            "ms_per_step": 1e3 * wall_m / args.steps,
            "p50_ms": 1e3 * float(np.percentile(lat_m, 50)),
            "p99_ms": 1e3 * float(np.percentile(lat_m, 99)),
            "cycles": [r[0] for r in res_m][0], "residual": res_m[0][1],
            "adm_mass": list(res_m[0][2]), "du_inf_vs_fp64": du,
            "vs_single_fp64": mixed_value / single_out["solves_per_s"],
            "roofline": roof_m,
        }
        assert du <= 1e-8, du

        pa = oa.problem_from_argv(BINARY)
        pa.precision = oa.PREC_MIXED
        pa.initial_guess = oa.GUESS_APPROX
        ab = Batch(oa, pa, args.batch, graphs)
        wall_a, lat_a, res_a = timed_batches(ab, args.warmup, args.steps)
        du_a = 0.0
        for ib in range(len(u64)):
            d = np.abs(ab.hiers[0].get_field(ib, oa.F_U) - u64[ib])[masks[ib]]
            du_a = max(du_a, float(d.max()))
        ab.close()
        approx_value = args.batch * args.steps / wall_a
        out["mixed_approx"] = {
            "solves_per_s": approx_value,
            "dtype": "f32 smoothing + f64 defect/update/transfers; u0 = analytic approximation",
            "ms_per_step": 1e3 * wall_a / args.steps,
            "p50_ms": 1e3 * float(np.percentile(lat_a, 50)),
            "p99_ms": 1e3 * float(np.percentile(lat_a, 99)),
            "cycles": res_a[0][0], "residual": res_a[0][1], "adm_mass": list(res_a[0][2]),
            "du_inf_vs_fp64": du_a, "vs_single_fp64": approx_value / single_out["solves_per_s"],
        }
        assert du_a <= 1e-8, du_a
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(BINARY, cycles)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
