"""Benchmark of the 3D channel-DNS time step on one MI355X.

Metric: grid-point steps per second (nx*ny*nz x full fractional steps / wall second) of the
production shape 128x128x129 (Lx = 4 pi, Ly = 2 pi, Re = 180, dt = 0.002, PI control on the
bulk velocity, 2/3 dealiasing, seeded stream-function noise of amplitude 0.05, exact LGL D1 --
the reference's thresholded D1 is diagonal at nz = 129 and the run goes non-finite within ~90
steps; both D1 choices cost the same on this path), FP64, state
resident in HBM, steps replayed from a captured HIP graph.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "grid-point RK4 steps/sec, FP64 plain vs mixed+batched; p50/p99 step latency"
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6

WORKLOADS = {
    "channel_128x128x129_fp64": dict(
        nx=128, ny=128, nz=129, xlen=4 * np.pi, ylen=2 * np.pi, re=180.0, dt=0.002, use_crank_nicolson=1,
        use_dealias_23_rule=1, flow_control_method=2, target_pressure_gradient=2.0 / 180.0,
        target_bulk_velocity=2.0 / 3.0, controller_gain=0.15, controller_update_freq=7, istart=0,
        u_centerline=1.0, enable_perturbations=1, perturbation_kind=1, seed=2024, noise_kmax=8,
        perturbation_amplitude=0.05, lgl_derivative=1),
    "channel_32x32x33_fp64": dict(
        nx=32, ny=32, nz=33, xlen=2 * np.pi, ylen=np.pi, re=180.0, dt=0.002, use_crank_nicolson=1,
        use_dealias_23_rule=1, flow_control_method=1, target_pressure_gradient=2.0 / 180.0,
        u_centerline=1.0, enable_perturbations=1, perturbation_kind=0, analytical_mode_x=1,
        analytical_mode_y=1, perturbation_amplitude=0.01),
}


def fold_macs(rows, cols):
    """Multiply-adds of an even/odd-folded rows x cols operator on one column: the even half
    ceil(rows/2) x ceil(cols/2) plus the odd half floor(rows/2) x floor(cols/2)."""
    return -(-rows // 2) * -(-cols // 2) + (rows // 2) * (cols // 2)


def kernel_work(kw):
    """Algorithmic HBM bytes and flops per launch of each step kernel (FP64).

    Spectral field: 16 B x (nx/2+1) ny nz; physical field: 8 B x nx ny nz.  Each kernel is
    charged one read of its inputs and one write of its outputs.  Flops count the useful
    multiply-adds (2 flops each) of the folded wall-normal contractions as computed: D1 on two
    fields; per component the Helmholtz forward [T | f T D2_I] (m x m on v, m x nz on u) and
    back B (m x m); the Poisson forward [Tf | Tf D1] and back [Bf ; D1 Bf] (four nz x nz).
    Transforms count none.
    """
    nx, ny, nz = kw["nx"], kw["ny"], kw["nz"]
    H2 = 2 * (nx // 2 + 1) * ny
    S = 8.0 * H2 * nz            # bytes of one spectral field
    P = 8.0 * nx * ny * nz       # bytes of one physical field
    m = nz - 2
    return {
        "zderiv_vort": dict(bytes=5 * S, flops=2.0 * 2 * fold_macs(nz, nz) * H2),
        "c2r_nl": dict(bytes=5 * S + 6 * P, flops=0.0),
        "r2c_cross": dict(bytes=6 * P + 3 * S, flops=0.0),
        "flow_control": dict(bytes=16.0 * nz, flops=0.0),
        "helmholtz": dict(bytes=9 * S, flops=2.0 * 3 * (2 * fold_macs(m, m) + fold_macs(m, nz)) * H2),
        "poisson_fused": dict(bytes=8 * S, flops=2.0 * 4 * fold_macs(nz, nz) * H2),
        "poisson_mode0": dict(bytes=0.0, flops=0.0),
    }


def kernel_fractions(prof, work):
    """Per step kernel: measured ms, algorithmic GB/s and FP64 TFLOP/s, and their fractions of
    the HBM and FP64 peaks (the larger is the kernel's roofline fraction)."""
    out = {}
    for name, ms in prof.items():
        if ms <= 0 or name not in work:
            continue
        wk = work[name]
        gbs = wk["bytes"] / (ms * 1e-3) / 1e9
        tfs = wk["flops"] / (ms * 1e-3) / 1e12
        out[name] = {"ms": ms, "algorithmic_gbs": gbs, "hbm_frac": gbs / PEAK_HBM_GBS, "fp64_tflops": tfs,
                     "fp64_frac": tfs / PEAK_FP64_TFLOPS, "frac": max(gbs / PEAK_HBM_GBS, tfs / PEAK_FP64_TFLOPS)}
    return out


def step_roofline(work, ms_per_step):
    """The fused step's own roof: the algorithmic bytes and FP64 flops of all its kernels (each
    charged one read of its inputs and one write of its outputs), the time each would take at
    peak, and the measured step against the larger of the two."""
    b = sum(w["bytes"] for w in work.values())
    f = sum(w["flops"] for w in work.values())
    t_hbm = b / (PEAK_HBM_GBS * 1e9) * 1e3
    t_fp64 = f / (PEAK_FP64_TFLOPS * 1e12) * 1e3
    roof_ms = max(t_hbm, t_fp64)
    return {"algorithmic_bytes": b, "fp64_flops": f, "hbm_bound_ms": t_hbm, "fp64_bound_ms": t_fp64,
            "achieved_gbs": b / (ms_per_step * 1e-3) / 1e9, "achieved_fp64_tflops": f / (ms_per_step * 1e-3) / 1e12,
            "frac": roof_ms / ms_per_step}


def profiled_traffic(profile, kernel):
    """HBM bytes per dispatch of a step kernel from a committed rocprof summary (FETCH_SIZE and
    WRITE_SIZE runs scaled by the calibration copy kernel); None when not profiled."""
    path = os.path.join(ROOT, "profiles", profile)
    if not os.path.exists(path):
        return None
    for line in open(path):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic" and parts[1] == kernel:
            return float(parts[2])
    return None


def cpu_info():
    """Host CPU model and physical core count (lscpu), and the threads this job may use."""
    import subprocess
    model, cores, sockets = "unknown", None, 1
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            v = v.strip()
            if k == "Model name":
                model = v
            elif k == "Core(s) per socket":
                cores = int(v)
            elif k == "Socket(s)":
                sockets = int(v)
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    physical = cores * sockets if cores else None
    allowed = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(x for x in (physical, allowed, env or None) if x)
    return {"model": model, "physical_cores": physical, "threads_available": threads}


def _set_threads(n):
    import ctypes
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(int(n))
    except OSError:
        pass


def _oracle_run(kw, steps, threads):
    """Per-step wall times of the oracle in the reference's as-written mode (per-mode matrix
    rebuild and LU every step, serial mode loops, OpenMP over planes/columns)."""
    from oracle import oracle as orc
    _set_threads(threads)
    o = orc.Oracle(refactor_every_step=1, **kw)
    o.init_fields()
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        o.step(1)
        times.append(time.perf_counter() - t0)
    return np.array(times)


def cpu_baseline(kw, steps, channel32_steps):
    """The reported CPU baseline: the oracle -- the C port of the reference
    algorithm as written -- on this host, the production workload for a few steps on all cores
    available, the 32x32x33 channel for channel32_steps steps on all cores and on one thread;
    wall-clock per step, set-up excluded."""
    info = cpu_info()
    thr = info["threads_available"]
    n = kw["nx"] * kw["ny"] * kw["nz"]
    t = _oracle_run(kw, steps, thr)
    out = {"value": n * steps / t.sum(), "unit": "grid-point-steps/s", "cores": thr, "kind": "port",
           "sample": f"{steps} step(s) of the same {kw['nx']}x{kw['ny']}x{kw['nz']} workload, FP64, as written "
                     f"(per-mode rebuild + LU every step), {thr} OpenMP threads, {t.sum():.1f} s wall",
           "cpu_model": info["model"], "physical_cores": info["physical_cores"],
           "step_s": {"p50": float(np.percentile(t, 50)), "p99": float(np.percentile(t, 99))}}
    if channel32_steps > 0:
        k1 = WORKLOADS["channel_32x32x33_fp64"]
        n1 = k1["nx"] * k1["ny"] * k1["nz"]
        for label, th in (("channel32_all_threads", thr), ("channel32_1_thread", 1)):
            t1 = _oracle_run(k1, channel32_steps, th)
            out[label] = {"value": n1 * channel32_steps / t1.sum(), "unit": "grid-point-steps/s", "threads": th,
                          "steps": channel32_steps, "steps_per_s": channel32_steps / t1.sum(),
                          "step_s": {"p50": float(np.percentile(t1, 50)), "p99": float(np.percentile(t1, 99))}}
        _set_threads(thr)
    return out


def ensemble_variant(dns, steps, warmup, precision, autotune=True):
    """The 64-member sweep (nx, ny in {32, 64, 96}, nz in {33, 65}, Re ~ U[100, 400], amplitude
    ~ logU[0.005, 0.1]) advanced as one ragged ensemble, one graph replay per step."""
    from f90_dns_amd.workloads import ensemble_members
    mem = ensemble_members(64, precision=precision)
    ens = dns.Ensemble(mem)
    ens.init_fields()
    tune = ens.autotune(steps_per_trial=5) if autotune else None
    ens.step(warmup)
    ens.synchronize()
    t0 = time.perf_counter()
    ens.step(steps)
    ens.synchronize()
    wall = time.perf_counter() - t0
    lat = ens.step_timing()
    pts = sum(ens.points(i) for i in range(len(mem)))
    finite = all(np.isfinite(ens.diagnostics(i)["energy"]) for i in range(0, len(mem), 7))
    out = {"value": pts * steps / wall, "unit": "grid-point-steps/s", "members": len(mem), "groups": ens.groups,
           "grid_points": pts, "steps": steps, "ms_per_step": wall / steps * 1e3,
           "step_latency_ms": {"p50": lat["p50_ms"], "p99": lat["p99_ms"]},
           "hbm_gib": ens.device_bytes / 2 ** 30, "finite": bool(finite),
           "precision": "mixed" if precision else "fp64",
           "tuning": tune["chosen"] if tune else None}
    ens.close()
    return out


def serving_variant(dns, njobs, rate, budget_gib, slots):
    """Served sweep: seeded Poisson arrivals of mixed-precision jobs (grids up to 192x192x129,
    nz in {65, 129}, 50-1000 steps), admitted by HBM budget, continuous batching per step."""
    from f90_dns_amd.workloads import served_jobs
    jobs = served_jobs(njobs, rate_per_s=rate)
    # set-up outside the timed region: the wall-normal eigenbases of each nz (a one-time
    # cost per process, like planning a transform)
    for nz in sorted({p["nz"] for _, p, _ in jobs}):
        dns.estimate_bytes(1, **[p for _, p, _ in jobs if p["nz"] == nz][0])
    srv = dns.Server(int(budget_gib * 2 ** 30), slots_per_group=slots)
    t0 = time.perf_counter()
    i = 0
    while True:
        now = time.perf_counter() - t0
        while i < len(jobs) and jobs[i][0] <= now:
            srv.submit(jobs[i][2], **jobs[i][1])
            i += 1
        running = srv.tick()
        if i == len(jobs) and srv.pending() == 0:
            break
        if running == 0 and srv.pending() == 0 and i < len(jobs):
            time.sleep(max(0.0, min(0.01, jobs[i][0] - (time.perf_counter() - t0))))
    srv.drain()
    wall = time.perf_counter() - t0
    st = srv.stats()
    turn = sorted(srv.times(j)["turnaround_s"] for j in range(1, len(jobs) + 1))
    out = {"value": st["grid_point_steps"] / wall, "unit": "grid-point-steps/s", "jobs": len(jobs),
           "arrival_rate_per_s": rate, "done": st["done"], "failed": st["failed"], "rejected": st["rejected"],
           "job_steps": st["job_steps"], "ticks": st["ticks"], "wall_s": wall,
           "device_busy_frac": st["device_ms"] * 1e-3 / wall,
           "job_step_latency_ms": {"p50": st["step_p50_ms"], "p99": st["step_p99_ms"]},
           "turnaround_s": {"p50": float(np.percentile(turn, 50)), "p99": float(np.percentile(turn, 99))},
           "hbm_budget_gib": budget_gib, "peak_hbm_gib": st["peak_bytes"] / 2 ** 30, "max_running": st["max_running"],
           "membership_changes": st["changes"], "change_host_ms": st["change_ms"],
           "graph_captures": st["captures"], "capture_host_ms": st["capture_ms"],
           "precision": "mixed"}
    srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="channel_128x128x129_fp64", choices=sorted(WORKLOADS))
    ap.add_argument("--no-graph", action="store_true", help="launch kernels eagerly instead of graph replay")
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--cpu-steps", type=int, default=2, help="oracle steps of the production workload")
    ap.add_argument("--cpu-channel32-steps", type=int, default=100, help="oracle steps of the 32x32x33 channel (0 = skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--variant-steps", type=int, default=200,
                    help="steps for the mixed-precision line reported beside the FP64 value (0 = skip)")
    ap.add_argument("--ensemble-steps", type=int, default=100, help="steps of the 64-member ensemble lines (0 = skip)")
    ap.add_argument("--serve-jobs", type=int, default=32, help="jobs of the served sweep (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=40.0, help="Poisson arrival rate of the served sweep (1/s)")
    ap.add_argument("--serve-budget-gib", type=float, default=24.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the ensemble/serving "
              "path would shard whole members across GPUs)", file=sys.stderr)
        sys.exit(2)

    import f90_dns_amd as dns
    kw = WORKLOADS[args.workload]
    sim = dns.Simulation(**kw)
    sim.init_fields()
    graph = not args.no_graph
    print(f"[bench] {args.workload}: warmup {args.warmup}", file=sys.stderr, flush=True)
    sim.step(args.warmup, graph=graph)
    sim.synchronize()
    print(f"[bench] timing {args.steps} steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    sim.step(args.steps, graph=graph)
    sim.synchronize()
    wall = time.perf_counter() - t0
    lat = sim.step_timing()
    npts = kw["nx"] * kw["ny"] * kw["nz"]
    value = npts * args.steps / wall
    diag = sim.diagnostics()
    if not np.isfinite(diag["energy"]):
        raise SystemExit("non-finite energy after the timed steps")

    # per-kernel device time (events between launches, same stream) for the roofline entry
    prof = sim.profile_step(args.profile_steps) if args.profile_steps > 0 else {}
    work = kernel_work(kw)
    roof = None
    if prof:
        name = max(prof, key=prof.get)
        ms = prof[name]
        wk = work[name]
        profile = "step_128x128x129_fp64_kernels.md" if args.workload == "channel_128x128x129_fp64" else None
        traffic = profiled_traffic(profile, name) if profile else None
        if wk["flops"] > 0:
            ach = wk["flops"] / (ms * 1e-3) / 1e12
            roof = {"kernel": name, "bound": "mfma", "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                    "frac": ach / PEAK_FP64_TFLOPS, "traffic": traffic, "ms": ms,
                    "hbm_gbs": wk["bytes"] / (ms * 1e-3) / 1e9}
        else:
            ach = wk["bytes"] / (ms * 1e-3) / 1e9
            roof = {"kernel": name, "bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                    "frac": ach / PEAK_HBM_GBS, "traffic": traffic, "ms": ms}
        roof["algorithmic_bytes"] = wk["bytes"]
        if traffic is not None:
            roof["traffic_source"] = "profiles/" + profile
        roof["kernels_ms"] = prof
        roof["kernels"] = kernel_fractions(prof, work)
    print("[bench] kernel ms: " + json.dumps(prof), file=sys.stderr, flush=True)

    variants = {}
    if args.variant_steps > 0:
        sim.close()
        # the 2D solver's 4-stage 1/irk convection on the same step (flagged variant)
        r = dns.Simulation(convection_scheme=1, **kw)
        r.init_fields()
        r.step(args.warmup, graph=graph)
        r.synchronize()
        t0 = time.perf_counter()
        r.step(args.variant_steps, graph=graph)
        r.synchronize()
        rw = time.perf_counter() - t0
        rl = r.step_timing()
        variants["rk4_convection_fp64"] = {
            "value": npts * args.variant_steps / rw, "unit": "grid-point-steps/s", "steps": args.variant_steps,
            "ms_per_step": rw / args.variant_steps * 1e3,
            "step_latency_ms": {"p50": rl["p50_ms"], "p99": rl["p99_ms"]},
            "scheme": "4-stage 1/irk convection (runge_kutta_convection, DNS_pressure_BC_2D.f90:931-1014) "
                      "in the 3D fractional step"}
        r.close()
        mixed = ((dns.DNS_PREC_MIXED, "mixed_precision",
                  "FP32 transforms, product, D1 output and S storage; FP64 D1 arithmetic and solves"),
                 (dns.DNS_PREC_MIXED_F32D, "mixed_precision_f32_contraction",
                  "as mixed_precision, and the vorticity's D1 contraction on FP32 MFMA (v_mfma_f32_16x16x4f32, "
                  "FP32 accumulation)"))
        for prec, name, what in mixed:
            m = dns.Simulation(precision=prec, **kw)
            m.init_fields()
            m.step(args.warmup, graph=graph)
            m.synchronize()
            t0 = time.perf_counter()
            m.step(args.variant_steps, graph=graph)
            m.synchronize()
            mw = time.perf_counter() - t0
            ml = m.step_timing()
            variants[name] = {
                "value": npts * args.variant_steps / mw, "unit": "grid-point-steps/s", "steps": args.variant_steps,
                "ms_per_step": mw / args.variant_steps * 1e3,
                "speedup_vs_fp64": (npts * args.variant_steps / mw) / value,
                "step_latency_ms": {"p50": ml["p50_ms"], "p99": ml["p99_ms"]},
                "kernels_ms": m.profile_step(args.profile_steps) if args.profile_steps > 0 else {},
                "precision": what}
            m.close()

    if args.ensemble_steps > 0:
        for prec, name in ((1, "ensemble64_mixed"), (0, "ensemble64_fp64")):
            print(f"[bench] {name}: {args.ensemble_steps} steps", file=sys.stderr, flush=True)
            v = ensemble_variant(dns, args.ensemble_steps, args.warmup, prec)
            v["vs_plain_fp64_128"] = v["value"] / value
            variants[name] = v
    if args.serve_jobs > 0:
        print(f"[bench] served sweep: {args.serve_jobs} jobs", file=sys.stderr, flush=True)
        v = serving_variant(dns, args.serve_jobs, args.serve_rate, args.serve_budget_gib, 4)
        v["vs_plain_fp64_128"] = v["value"] / value
        variants["served_sweep_mixed"] = v

    cpu = None
    if not args.no_cpu_baseline and args.cpu_steps > 0:
        print(f"[bench] cpu baseline: {args.cpu_steps} oracle step(s), 32x32x33 x {args.cpu_channel32_steps}",
              file=sys.stderr, flush=True)
        cpu = cpu_baseline(kw, args.cpu_steps, args.cpu_channel32_steps)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "grid-point-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded stream-function noise initial field)",
        "config": {"workload": args.workload, "nx": kw["nx"], "ny": kw["ny"], "nz": kw["nz"],
                   "precision": "fp64", "graph": graph, "parallelism": "single GPU",
                   "step": "one full fractional step of DNS_pressure_BC_3D.f90:314-326 (explicit-Euler "
                           "convection, Crank-Nicolson Helmholtz, Poisson projection); the 3D reference has no "
                           "RK4 -- the 2D solver's 4-stage convection is variants.rk4_convection_fp64"},
        "step_latency_ms": {"p50": lat["p50_ms"], "p99": lat["p99_ms"], "min": lat["min_ms"], "mean": lat["mean_ms"]},
        "roofline": roof,
        "step_roofline": step_roofline(work, wall / args.steps * 1e3),
        "cpu_baseline": cpu,
        "variants": variants,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
