"""The named workloads of the puncture path and a runner that solves one in
plain FP64 and in mixed precision on one GPU and reports time, cycles,
residual, ADM masses and max |u_mixed - u_FP64| over every determined point.

    python -m ollipticid_amd.workloads binary_hires     # N=129, O8, 9 levels
    python -m ollipticid_amd.workloads scalar_shell     # 3 punctures + scalar field

Unspecified parameters of the scalar-shell workload are fixed here and
recorded in the output: L = 256, nu2 = 20, tol = 1e-10, seed 2026 for the
puncture positions (|x| < 6, separation >= 2).  Its scalar field
phi = A exp(-(r - r0)^2 / sigma^2) with A = 0.01, r0 = 10, sigma = 2 is the
reference's `-psc 2` profile phi0 (a2^2/(1+a2^2)) exp(-(r-r0)^2/ps_sigma),
i.e. -phi0 0.02 -a2 1 -sigma 4 -r0 10.
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

PUNC = ["-eq", "Punctures", "-B", "Robin", "-BA", "0", "-Bn", "1"]
BINARY = ["-NumBH", "2", "-bhmp1", "0.487209", "-bhy1", "5.5",
          "-bhpx1", "-0.0901099", "-bhpy1", "-0.000703975",
          "-bhmp2", "0.487209", "-bhy2", "-5.5",
          "-bhpx2", "0.0901099", "-bhpy2", "0.000703975"]


def scalar_shell_argv(seed=2026):
    rng = np.random.default_rng(seed)
    pos = []
    while len(pos) < 3:
        x = rng.uniform(-6.0, 6.0, 3)
        if np.linalg.norm(x) < 6.0 and all(np.linalg.norm(x - y) >= 2.0 for y in pos):
            pos.append(x)
    argv = ["-eq", "Punctures_scalar_field", "-B", "Robin", "-BA", "0", "-Bn", "1",
            "-psc", "2", "-phi0", "0.02", "-a2", "1", "-r0", "10", "-sigma", "4", "-NumBH", "3"]
    for i, (m, x) in enumerate(zip((0.4, 0.3, 0.3), pos), start=1):
        argv += ["-bhmp%d" % i, "%.17g" % m, "-bhx%d" % i, "%.17g" % x[0],
                 "-bhy%d" % i, "%.17g" % x[1], "-bhz%d" % i, "%.17g" % x[2]]
    return argv + ["-L", "256", "-N", "161", "-l", "10", "-O", "6", "-nu2", "20", "-s", "1e-10"]


WORKLOADS = {
    "single_boosted": PUNC + ["-NumBH", "1", "-bhmp1", "1.0", "-bhpx1", "0.2", "-L", "256",
                              "-N", "33", "-l", "4", "-O", "2", "-nu2", "10", "-s", "1e-10"],
    "binary": PUNC + BINARY + ["-L", "1000", "-N", "65", "-l", "9", "-O", "4", "-nu2", "20",
                               "-s", "1e-11"],
    "binary_hires": PUNC + BINARY + ["-L", "1000", "-N", "129", "-l", "9", "-O", "8",
                                     "-nu2", "20", "-s", "1e-11"],
    "scalar_shell": scalar_shell_argv(),
}


def determined(h):
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


def run(name):
    import ollipticid_amd as oa

    argv = WORKLOADS[name]
    out = {"workload": name, "argv": " ".join(argv)}
    fields = {}
    for prec in ("fp64", "mixed"):
        p = oa.problem_from_argv(argv)
        p.precision = oa.PREC_MIXED if prec == "mixed" else oa.PREC_FP64
        with oa.Hierarchy(p) as h:
            t0 = time.perf_counter()
            cycles, hist, res = h.solve()
            ms_first = 1e3 * (time.perf_counter() - t0)
            h.reset()
            t0 = time.perf_counter()
            cycles, hist, res = h.solve()
            ms = 1e3 * (time.perf_counter() - t0)
            adm, _ = h.adm_mass()
            fields[prec] = [h.get_field(ib, oa.F_U) for ib in range(h.nbox)]
            masks = determined(h)
            out[prec] = {"cycles": cycles, "residual": res, "ms_per_solve": ms,
                         "ms_first_solve": ms_first, "adm_mass": list(adm),
                         "coarse_sweeps": h.last_solve_coarse_sweeps()}
            out["boxes"] = h.nbox
            out["levels"] = h.levels
            out["points"] = int(sum(np.prod(b["n"]) for b in h.boxes()))
            out["hbm_bytes"] = oa.problem_footprint(p)
            print("%s %s: %d cycles, residual %.3g, %.1f ms" % (name, prec, cycles, res, ms),
                  file=sys.stderr, flush=True)
    out["du_inf_mixed_vs_fp64"] = max(float(np.abs(a - b)[m].max())
                                      for a, b, m in zip(fields["mixed"], fields["fp64"], masks))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("workload", choices=sorted(WORKLOADS))
    args = ap.parse_args()
    print(json.dumps(run(args.workload)), flush=True)


if __name__ == "__main__":
    main()
