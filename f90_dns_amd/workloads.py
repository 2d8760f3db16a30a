"""Seeded workload generators for the ensemble and serving paths.

Both are parameter sweeps of the 3D channel run of DNS_pressure_BC_3D.f90 (grid, Re,
perturbation amplitude), with the project's seeded stream-function noise as initial
perturbation (the reference's random perturbation is clock-seeded,
perturbation_module.f90:79-84, so it cannot be reproduced).  Members use the exact LGL D1
(lgl_derivative = 1): with the reference's thresholded D1 the mode-(0,0) pressure matrix is
singular at nz = 65 (lgl_module.f90:211-263 zeroes every off-diagonal entry there).
"""
import numpy as np

DNS_PREC_MIXED = 1

BASE = dict(xlen=2 * np.pi, ylen=np.pi, dt=0.002, use_crank_nicolson=1, use_dealias_23_rule=1,
            flow_control_method=1, u_centerline=1.0, enable_perturbations=1, perturbation_kind=1,
            noise_kmax=4, lgl_derivative=1)


def ensemble_members(n=64, seed=4, precision=DNS_PREC_MIXED, grids=(32, 64, 96), nzs=(33, 65)):
    """n members: nx, ny drawn from grids, nz from nzs, Re ~ U[100, 400], amplitude ~
    logU[0.005, 0.1], constant forcing G = 2/Re (laminar centreline velocity 1), one noise
    seed per member."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        re = float(rng.uniform(100.0, 400.0))
        amp = float(np.exp(rng.uniform(np.log(0.005), np.log(0.1))))
        m = dict(BASE, nx=int(rng.choice(grids)), ny=int(rng.choice(grids)), nz=int(rng.choice(nzs)), re=re,
                 target_pressure_gradient=2.0 / re, perturbation_amplitude=amp, seed=1000 + i,
                 precision=precision)
        out.append(m)
    return out


def served_jobs(njobs=48, seed=5, rate_per_s=40.0, precision=DNS_PREC_MIXED, grids=(64, 96, 128, 192),
                nzs=(65, 129), steps=(50, 1000)):
    """Poisson arrivals at rate_per_s (seeded): list of (arrival_s, params, nsteps) with grids up
    to 192 x 192 x 129, Re ~ U[100, 400], amplitude ~ U[0.005, 0.1], 50-1000 steps each."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for i in range(njobs):
        t += float(rng.exponential(1.0 / rate_per_s))
        re = float(rng.uniform(100.0, 400.0))
        p = dict(BASE, nx=int(rng.choice(grids)), ny=int(rng.choice(grids)), nz=int(rng.choice(nzs)), re=re,
                 target_pressure_gradient=2.0 / re, perturbation_amplitude=float(rng.uniform(0.005, 0.1)),
                 seed=5000 + i, precision=precision, xlen=4 * np.pi, ylen=2 * np.pi)
        jobs.append((t, p, int(rng.integers(steps[0], steps[1] + 1))))
    return jobs
