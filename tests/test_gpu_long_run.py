"""Long runs of the 128 x 128 x 129 production workload (the bench's: PI control of the bulk
velocity, seeded noise of amplitude 0.05), 10000 steps in FP64 and in mixed precision.  The
perturbations decay in this laminar regime, so the two runs must stay together (energy, bulk
velocity and forcing within 1e-7; measured: nine digits), finite, with the controller holding U_b near its target and the
divergence left by the collocation projection decaying."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KW = dict(nx=128, ny=128, nz=129, xlen=4 * np.pi, ylen=2 * np.pi, re=180.0, dt=0.002, use_crank_nicolson=1,
          use_dealias_23_rule=1, flow_control_method=2, target_pressure_gradient=2.0 / 180.0,
          target_bulk_velocity=2.0 / 3.0, controller_gain=0.15, controller_update_freq=7, istart=0,
          u_centerline=1.0, enable_perturbations=1, perturbation_kind=1, seed=2024, noise_kmax=8,
          perturbation_amplitude=0.05, lgl_derivative=1)


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dns():
    import f90_dns_amd
    f90_dns_amd.lib()
    return f90_dns_amd


def _run(dns, precision):
    s = dns.Simulation(precision=precision, **KW)
    s.init_fields()
    trace = []
    for _ in range(5):
        s.step(2000, graph=True)
        d = s.diagnostics()
        d["div"] = s.max_divergence()
        trace.append(d)
        _say(f"precision {precision} step {d['step']}: E {d['energy']:.9e} Ub {d['bulk_velocity']:.9f} "
             f"G {d['pressure_gradient']:.9e} div {d['div']:.3e}")
    s.close()
    return trace


def test_10000_steps_fp64_and_mixed(dns):
    t64 = _run(dns, dns.DNS_PREC_FP64)
    t32 = _run(dns, dns.DNS_PREC_MIXED)
    for a, b in zip(t64, t32):
        for d in (a, b):
            assert all(np.isfinite(d[k]) for k in ("energy", "bulk_velocity", "pressure_gradient", "umax", "wmax"))
            assert abs(d["bulk_velocity"] - 2.0 / 3.0) < 0.05
            assert d["div"] < 2e-3
        assert abs(b["energy"] - a["energy"]) < 1e-7 * a["energy"]
        assert abs(b["bulk_velocity"] - a["bulk_velocity"]) < 1e-7
        assert abs(b["pressure_gradient"] - a["pressure_gradient"]) < 1e-7 * abs(a["pressure_gradient"])
    # the perturbation decays: so does the divergence the projection leaves next to the walls
    assert t64[-1]["div"] < 0.5 * t64[0]["div"]
    assert t64[-1]["vmax"] < t64[0]["vmax"]
