"""Profiling driver: a few eager steps of a named workload, then a plain copy kernel of known
size (the calibration of a profiler's HBM byte counters).  Meant to run under rocprofv3, e.g.

    rocprofv3 --kernel-trace --stats -d OUT -o kt -- python3 -m f90_dns_amd.profile_step
    rocprofv3 --pmc FETCH_SIZE -d OUT -o pmc -- python3 -m f90_dns_amd.profile_step

Every kernel of the step is launched eagerly (one dispatch per kernel and step) so that the
profiler sees them one by one.
"""
import argparse
import ctypes
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

COPY_DOUBLES = 64 << 20   # 512 MiB read + 512 MiB written per copy launch


def main():
    import bench
    import f90_dns_amd as dns
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="channel_128x128x129_fp64", choices=sorted(bench.WORKLOADS) + ["ensemble64"])
    ap.add_argument("--precision", type=int, default=0)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=3)
    args = ap.parse_args()
    if args.workload == "ensemble64":
        from f90_dns_amd.workloads import ensemble_members
        sim = dns.Ensemble(ensemble_members(64, precision=args.precision))
        sim.init_fields()
        sim.set_tuning(lanes=1)
        step = lambda n: sim.step(n, graph=False)
    else:
        sim = dns.Simulation(precision=args.precision, **bench.WORKLOADS[args.workload])
        sim.init_fields()
        step = lambda n: sim.step(n, graph=False)
    step(args.warmup)
    sim.synchronize()
    print(f"[profile] {args.workload}: warm-up done", flush=True)
    step(args.steps)
    sim.synchronize()
    print(f"[profile] {args.steps} steps done", flush=True)
    ms = ctypes.c_double()
    rc = dns.lib().dns_debug_copy(COPY_DOUBLES, 3, ctypes.byref(ms))
    if rc != 0:
        raise SystemExit(f"calibration copy failed ({rc})")
    print(f"[profile] calibration copy: {COPY_DOUBLES * 8 / 2**20:.0f} MiB each way, {ms.value:.3f} ms", flush=True)


if __name__ == "__main__":
    main()
