"""The 2D solver's numerics on the 3D step, behind flags (parity unpinned: the reference has no
3D trace of them; the checker is the oracle's restatement of runge_kutta_convection,
DNS_pressure_BC_2D.f90:931-1014, and jacgs, :2361-2476, applied to the 3D step)."""
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_util import CHANNEL32, W_SELF_FP64, W_SELF_MIXED, check_parity, field_errors, rel_l2

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dns():
    import f90_dns_amd
    f90_dns_amd.lib()
    return f90_dns_amd


def _pair(dns, **kw):
    o = orc.Oracle(**{k: v for k, v in kw.items() if k != "precision"})
    o.init_fields()
    g = dns.Simulation(**kw)
    g.set_fields(*o.get_fields())
    return o, g


@pytest.mark.parametrize("kw", [dict(CHANNEL32),
                                dict(CHANNEL32, nx=48, ny=36, nz=129, lgl_derivative=1, flow_control_method=2,
                                     perturbation_kind=1, seed=3, noise_kmax=5, perturbation_amplitude=0.05),
                                dict(CHANNEL32, nx=48, ny=36, nz=34, lgl_derivative=1, flow_control_method=2,
                                     perturbation_kind=1, seed=17, noise_kmax=5, perturbation_amplitude=0.05),
                                dict(CHANNEL32, nx=48, ny=36, nz=140, lgl_derivative=1, flow_control_method=2,
                                     perturbation_kind=1, seed=18, noise_kmax=5, perturbation_amplitude=0.05)],
                         ids=["channel32", "48x36x129", "even48x36x34", "pairs48x36x140"])
def test_rk4_convection_against_oracle(dns, kw):
    _say(f"rk4 convection {kw['nx']}x{kw['ny']}x{kw['nz']}")
    o, g = _pair(dns, convection_scheme=1, **kw)
    o.step(20)
    g.step(20)
    e = field_errors(g.get_fields(), o.get_fields())
    _say(f"  after 20 steps {e}")
    check_parity(e, 1e-10, W_SELF_FP64)
    d_o, d_g = o.diagnostics(), g.diagnostics()
    assert abs(d_o["energy"] - d_g["energy"]) < 1e-12 * d_o["energy"]


def test_rk4_stage_api_graph_and_fused_agree(dns):
    kw = dict(CHANNEL32, convection_scheme=1)
    _, a = _pair(dns, **kw)
    _, b = _pair(dns, **kw)
    _, c = _pair(dns, **kw)
    for _ in range(4):
        a.step(1)
        b.begin_step(); b.compute_source_terms(); b.viscous_step(); b.solve_pressure()
        b.pressure_correction(); b.apply_boundary_conditions()
    c.step(2, graph=True)
    c.step(2, graph=True)
    for x, y, z in zip(a.get_fields(), b.get_fields(), c.get_fields()):
        assert rel_l2(y, x) < 1e-14
        np.testing.assert_array_equal(z, x)


@pytest.mark.parametrize("nz", [34, 66, 130])
def test_stage_api_matches_oracle_at_even_nz(dns, nz):
    """The reference's stage sequence (every plane transformed, the Poisson solve and the
    correction as separate stages) at even nz, where the folds have no centre row, including
    two row-block pairs per wave (nz = 130): 3 steps against the oracle and the fused step."""
    kw = dict(CHANNEL32, nx=48, ny=36, nz=nz, lgl_derivative=1, flow_control_method=2, perturbation_kind=1, seed=19,
              noise_kmax=5, perturbation_amplitude=0.05)
    o, a = _pair(dns, **kw)
    _, b = _pair(dns, **kw)
    o.step(3)
    a.step(3)
    for _ in range(3):
        b.begin_step(); b.compute_source_terms(); b.viscous_step(); b.solve_pressure()
        b.pressure_correction(); b.apply_boundary_conditions()
    e = field_errors(b.get_fields(), o.get_fields())
    _say(f"stage api 48x36x{nz} after 3 steps {e}")
    check_parity(e, 1e-10, W_SELF_FP64)
    for x, y in zip(a.get_fields(), b.get_fields()):
        assert rel_l2(y, x) < 1e-13


def test_rk4_convection_mixed_precision(dns):
    o, g = _pair(dns, convection_scheme=1, precision=dns.DNS_PREC_MIXED, **CHANNEL32)
    o.step(100)
    g.step(100)
    e = field_errors(g.get_fields(), o.get_fields())
    _say(f"rk4 mixed after 100 steps {e}")
    check_parity(e, 1e-6, W_SELF_MIXED)


@pytest.mark.parametrize("tol,bound", [(1e-12, 1e-10), (1e-6, 1e-7)])
def test_jacgs_pressure_against_oracle(dns, tol, bound):
    """jacgs per real column on the GPU against the oracle's jacgs; with a tight cgstol both
    reach the direct solution, with the reference's default 1e-6 they stop (or skip modes whose
    RHS is below it) at rounding-dependent points, so they agree to that looser level."""
    kw = dict(CHANNEL32, flow_control_method=2, perturbation_kind=1, seed=12, noise_kmax=4,
              perturbation_amplitude=0.05, pressure_solver=1, cgstol=tol)
    o, g = _pair(dns, **kw)
    for _ in range(5):
        o.step(2)
        g.step(2)
        it_o, nf_o = o.jacgs_stats()
        it_g, nf_g = g.jacgs_stats()
        _say(f"  jacgs tol {tol}: iterations gpu {it_g} oracle {it_o}, non-converged {nf_g} {nf_o}")
        assert nf_g == 0 and nf_o == 0
        assert abs(it_g - it_o) <= 10 + 0.2 * it_o
    e = field_errors(g.get_fields(), o.get_fields())
    _say(f"  jacgs tol {tol} after 10 steps {e}")
    check_parity(e, bound, max(W_SELF_FP64, 1e3 * bound))


def test_jacgs_and_rk4_together_and_in_a_batch(dns):
    base = dict(CHANNEL32, convection_scheme=1, pressure_solver=1, cgstol=1e-12)
    mem = [dict(re=180.0), dict(re=260.0, flow_control_method=2, perturbation_kind=1, seed=4,
                                perturbation_amplitude=0.04)]
    batch = dns.Simulation(members=mem, **base)
    batch.init_fields()
    batch.step(6)
    for b, m in enumerate(mem):
        o = orc.Oracle(**{**base, **m})
        o.init_fields()
        o.step(6)
        e = field_errors(batch.get_fields(member=b), o.get_fields())
        _say(f"  rk4 + jacgs member {b}: {e}")
        check_parity(e, 1e-10, W_SELF_FP64)


def test_variants_rejected_where_not_built(dns):
    with pytest.raises(dns.DnsError):
        dns.Simulation(**dict(CHANNEL32, pressure_solver=1, use_weak_form_divergence=1))
    with pytest.raises(dns.DnsError):
        dns.Simulation(**dict(CHANNEL32, pressure_solver=1, precision=dns.DNS_PREC_MIXED))
    srv = dns.Server(hbm_budget_bytes=1 << 30)
    with pytest.raises(dns.DnsError):
        srv.submit(5, **dict(CHANNEL32, convection_scheme=1))
    srv.close()
