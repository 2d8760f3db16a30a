"""Write the compact oracle fixtures (tests/golden/oracle_<case>.npz) of the long runs.

Runs the CPU FP64 oracle (cached per-mode factors, bitwise the refactor-every-step mode) on
each case from its initial field, with a twin whose plane transforms use the reversed Stockham
factor order (fft_variant = 1) to record each field's rounding floor, and stores at each
snapshot step the record of tests/oracle_fixtures.py.  Cases:

- channel128: 128 x 128 x 129, Lx = 4 pi, Ly = 2 pi, Re = 180, dt = 0.002, constant bulk velocity
  (PI, U_b = 2/3, gain 0.15), seeded stream-function noise |kx|, |ky| <= 8, amplitude 0.05,
  exact LGL D1; snapshots after 100 and 500 steps.
- served_192x192x129, served_128x96x65: two jobs of the served-sweep test mix (Lx = 4 pi,
  Ly = 2 pi, constant forcing 2/Re, seeded noise |k| <= 4).
- ens_<i>: members 15, 35, 41 and 60 of the 64-member ensemble (f90_dns_amd.workloads,
  seed 4), 500 steps.

Usage: python tests/golden/make_oracle_fixtures.py [case ...]   (about 30 min on 8 cores)
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402
from tests import oracle_fixtures as fx  # noqa: E402

CHANNEL128 = dict(nx=128, ny=128, nz=129, xlen=4 * np.pi, ylen=2 * np.pi, re=180.0, dt=0.002, use_crank_nicolson=1,
               use_dealias_23_rule=1, flow_control_method=2, target_pressure_gradient=2.0 / 180.0,
               target_bulk_velocity=2.0 / 3.0, controller_gain=0.15, controller_update_freq=7, istart=0,
               u_centerline=1.0, enable_perturbations=1, perturbation_kind=1, seed=2024, noise_kmax=8,
               perturbation_amplitude=0.05, lgl_derivative=1)


def served(nx, ny, nz, re, amp, seed):
    from f90_dns_amd.workloads import BASE
    return dict(BASE, nx=nx, ny=ny, nz=nz, re=re, target_pressure_gradient=2.0 / re, perturbation_amplitude=amp,
                seed=seed, xlen=4 * np.pi, ylen=2 * np.pi)


def ensemble_member(i):
    from f90_dns_amd.workloads import ensemble_members
    return {k: v for k, v in ensemble_members()[i].items() if k != "precision"}


CASES = {
    "channel128": (CHANNEL128, [100, 500]),
    "served_192x192x129": (served(192, 192, 129, 250.0, 0.07, 5101), [60]),
    "served_128x96x65": (served(128, 96, 65, 320.0, 0.02, 5102), [200]),
}
for _i in (15, 35, 41, 60):
    CASES[f"ens_{_i}"] = (None, [500])


def params_of(name):
    p, steps = CASES[name]
    if p is None:
        p = ensemble_member(int(name.split("_")[1]))
    return dict(p), steps


def diag_row(o):
    d = o.diagnostics()
    return {k: d[k] for k in ("step", "umax", "vmax", "wmax", "energy", "pressure_gradient", "bulk_velocity",
                               "avg_pressure_gradient", "bulk_error_integral")}


def run(name):
    p, steps = params_of(name)
    t0 = time.time()
    a = orc.Oracle(**p)
    a.init_fields()
    b = orc.Oracle(fft_variant=1, **p)
    b.init_fields()
    plane = p["nz"] // 4
    trace, snaps, floors = [], {}, {}
    done = 0
    for s in steps:
        while done < s:
            n = min(10, s - done)
            a.step(n); b.step(n)
            done += n
            trace.append(diag_row(a))
            print(f"{name}: step {done} ({time.time() - t0:.0f} s)", flush=True)
        fa, fb = a.get_fields(), b.get_fields()
        snaps[s] = fx.summarize(fa, plane)
        floors[str(s)] = {c: float(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))
                          for c, x, y in zip(fx.FIELDS, fb, fa)}
        floors[str(s)]["max_divergence"] = a.max_divergence()
    meta = {"case": name, "params": {k: (float(v) if isinstance(v, float) else v) for k, v in p.items()},
            "steps": steps, "plane": plane, "trace": trace, "floor": floors,
            "sketch": {"rows": fx.SKETCH_ROWS, "seed": fx.SKETCH_SEED, "chunk": fx.SKETCH_CHUNK}}
    fx.save(name, meta, snaps)
    print(f"{name}: wrote {fx.path(name)} ({os.path.getsize(fx.path(name)) // 1024} KB) in {time.time() - t0:.0f} s; "
          f"floors {floors}", flush=True)


if __name__ == "__main__":
    for name in sys.argv[1:] or list(CASES):
        run(name)
