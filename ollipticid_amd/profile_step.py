"""One batched solve of a named workload, issued launch by launch so that a
profiler can attribute kernel durations and hardware counters to each
dispatch, plus an optional device-to-device copy of known size that
calibrates the HBM byte counters.

    rocprofv3 --pmc FETCH_SIZE --kernel-include-regex 'k_smooth|copy' \\
        -- python3 -m ollipticid_amd.profile_step --batch 16 --cycles 2 --copy-mib 512
    rocprofv3 --kernel-trace --stats \\
        -- python3 -m ollipticid_amd.profile_step --workload scalar_shell --batch 1 --cycles 0

Workloads are those of ollipticid_amd.workloads (default: the equal-mass
binary, N=65, 9 levels, 4th order).  Prints one JSON line: the workload, the
number of V-cycles run and the smoother launches, device time and
algorithmic bytes over the run.
"""
import argparse
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    from ollipticid_amd.workloads import WORKLOADS

    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="binary")
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--cycles", type=int, default=2, help="V-cycles per request (0: to convergence)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also copy this many MiB device to device (counter calibration)")
    ap.add_argument("--fused", action="store_true",
                    help="fused double-buffered FP64 sweeps instead of colour by colour")
    ap.add_argument("--fused-shape", default=None,
                    help="ty,tz,threads of the fused sweeps' tiles (with --fused)")
    args = ap.parse_args()

    import ollipticid_amd as oa
    from ollipticid_amd import native

    p = oa.problem_from_argv(WORKLOADS[args.workload])
    native.set_fused_sweeps(args.fused)
    if args.fused_shape:
        native.set_fused_shape(p.order, *[int(v) for v in args.fused_shape.split(",")])
    p.precision = oa.PREC_MIXED if args.precision == "mixed" else oa.PREC_FP64
    if args.cycles:
        p.max_cycles = args.cycles
    e = oa.Engine(graphs=False)
    for t in range(args.batch):
        assert e.admit(p, t)
    e.set_kernel_timing(True)
    cycles = 0
    while e.step():
        cycles += 1
    cycles += 1
    launches, ms, nbytes = e.kernel_timing()
    done = e.pop_done()
    e.close()
    out = {"workload": args.workload, "argv": " ".join(WORKLOADS[args.workload]),
           "batch": args.batch, "precision": args.precision, "fused": args.fused,
           "fused_shape": native.fused_shape(p.order) if args.fused else None,
           "cycles": cycles,
           "requests_done": len(done), "smoother_launches": launches, "smoother_ms": ms,
           "smoother_bytes": nbytes}
    if args.copy_mib:
        # the HIP runtime's device-to-device copy kernel on buffers of known size
        import ctypes as C

        hip = C.CDLL("libamdhip64.so")
        n = args.copy_mib << 20
        src, dst = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(src), C.c_size_t(n)) == 0
        assert hip.hipMalloc(C.byref(dst), C.c_size_t(n)) == 0
        assert hip.hipMemset(src, 1, C.c_size_t(n)) == 0
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(dst, src, C.c_size_t(n), 3) == 0  # hipMemcpyDeviceToDevice
        assert hip.hipDeviceSynchronize() == 0
        hip.hipFree(src)
        hip.hipFree(dst)
        out["copy_bytes"] = n
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
