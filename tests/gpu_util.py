"""Helpers shared by the GPU parity tests (oracle = CPU FP64 restatement)."""
import numpy as np

from oracle import oracle as orc


def rel_l2(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def field_errors(got, ref):
    """Per-field parity errors for (u, v, w, p): relative L2 for u, v, p; for w, the L2 error
    relative to the velocity-field norm ||(u, v, w)|| -- w is ~1e-6 of u in these channel runs,
    so its self-relative error sits on the rounding floor of the FFT round trips (~2e-10 at
    the 32x32x33 channel); that self-relative value is returned as 'w_self'."""
    u, v, w = ref[:3]
    p = ref[3] if len(ref) > 3 else None
    vel = np.sqrt(np.linalg.norm(u) ** 2 + np.linalg.norm(v) ** 2 + np.linalg.norm(w) ** 2)
    e = {"u": rel_l2(got[0], u), "v": rel_l2(got[1], v),
         "w": np.linalg.norm(got[2] - w) / vel, "w_self": rel_l2(got[2], w)}
    if len(got) > 3 and len(ref) > 3:
        e["p"] = rel_l2(got[3], p)
    return e


def max_err(e):
    return max(v for k, v in e.items() if k != "w_self")


# Bounds on w's error relative to its own norm.  w is ~1e-6 of u in these runs, so rounding
# moves it at ~1e-10 of its own norm under any reordering of the arithmetic: the oracle against
# its own twin with other plane-transform factorisations disagrees by 1.2e-10 in w after 100
# 32x32x33 steps (tests/test_oracle_pins.py::test_w_rounding_floor_of_the_oracle_itself) while
# u, v, p agree to 1e-13.  FP64 paths are held to W_SELF_FP64 (a few times the largest such
# floor seen, 2e-9 at 128^2 x 129 after 3 steps, where w is younger and smaller still); the
# long runs compare against the floor measured for that very run.  Mixed precision rounds the
# product in FP32 (6e-8 of |u x omega|), which is ~1e-1 of w's own source term: W_SELF_MIXED
# records that known miss against the 1e-6 target (u, v, p and w on the velocity scale meet it).
W_SELF_FP64 = 5e-9
W_SELF_MIXED = 1e-3


def check_parity(e, tol, w_self_tol):
    """u, v, p relative L2, w on the velocity scale below tol; w's own relative L2 below
    w_self_tol."""
    assert max_err(e) < tol, e
    assert e["w_self"] < w_self_tol, e


# the 32x32x33 channel (32x32x33, Lx=2pi, Ly=pi, Re=180, dt=0.002, G=2/Re, analytical mx=my=1 amp 0.01, dealias)
CHANNEL32 = dict(nx=32, ny=32, nz=33, xlen=2 * np.pi, ylen=np.pi, re=180.0, dt=0.002, use_crank_nicolson=1,
               use_dealias_23_rule=1, flow_control_method=1, target_pressure_gradient=2.0 / 180.0,
               target_bulk_velocity=2.0 / 3.0, controller_gain=0.15, controller_update_freq=7, istart=0,
               u_centerline=1.0, enable_perturbations=1, perturbation_kind=0, analytical_mode_x=1,
               analytical_mode_y=1, perturbation_amplitude=0.01, seed=0, noise_kmax=8)


def pair(**kw):
    """An oracle and a GPU simulation on identical initial fields (the oracle's)."""
    import f90_dns_amd as dns
    o = orc.Oracle(**kw)
    o.init_fields()
    g = dns.Simulation(**kw)
    u, v, w, p = o.get_fields()
    g.set_fields(u, v, w, p)
    return o, g
