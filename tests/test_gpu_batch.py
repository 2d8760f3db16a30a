"""Batched contexts: B members of one shape advanced by the same launches.

A member's result must not depend on the company it keeps: every member is computed by the
same arithmetic as a context of one, so the fields and diagnostics of a member in a batch are
bitwise those of the same run alone.  Members are also checked against the FP64 oracle."""
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_util import CHANNEL32, W_SELF_FP64, W_SELF_MIXED, check_parity, field_errors, max_err

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dns():
    import f90_dns_amd
    f90_dns_amd.lib()
    return f90_dns_amd


def _members(nz):
    """Four members of one shape: different Re, controller, perturbation kind, seed, amplitude."""
    return [
        dict(re=180.0, flow_control_method=1, perturbation_kind=0, perturbation_amplitude=0.01),
        dict(re=120.0, flow_control_method=2, perturbation_kind=1, seed=11, perturbation_amplitude=0.05),
        dict(re=390.0, flow_control_method=1, target_pressure_gradient=2.0 / 390.0, perturbation_kind=1, seed=12,
             perturbation_amplitude=0.006),
        dict(re=260.0, flow_control_method=2, perturbation_kind=1, seed=13, perturbation_amplitude=0.09),
    ]


SHAPES = [
    dict(nx=32, ny=32, nz=33),
    dict(nx=64, ny=96, nz=65, xlen=4 * np.pi, ylen=2 * np.pi, lgl_derivative=1),
    # even nz (no centre node) and two row-block pairs per wave
    dict(nx=48, ny=36, nz=66, lgl_derivative=1),
    dict(nx=48, ny=36, nz=130, lgl_derivative=1),
]


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("shape", SHAPES, ids=["32x32x33", "64x96x65", "48x36x66", "48x36x130"])
def test_member_alone_equals_member_in_batch(dns, shape, precision):
    base = {**CHANNEL32, **shape, "precision": precision}
    mem = _members(shape["nz"])
    _say(f"batch of {len(mem)} at {shape}, precision {precision}")
    batch = dns.Simulation(members=mem, **base)
    batch.init_fields()
    batch.step(25, graph=True)
    batch.synchronize()
    for b, m in enumerate(mem):
        one = dns.Simulation(**{**base, **m})
        one.init_fields()
        one.step(25, graph=True)
        got, ref = batch.get_fields(member=b), one.get_fields()
        for f in range(4):
            assert np.array_equal(got[f], ref[f]), (b, f, np.max(np.abs(got[f] - ref[f])))
        db, d1 = batch.diagnostics(member=b), one.diagnostics()
        assert db == d1, (b, db, d1)
        assert np.isfinite(db["energy"])
        one.close()
    batch.close()


def test_batch_eager_equals_graph(dns):
    base = {**CHANNEL32}
    mem = _members(33)
    a = dns.Simulation(members=mem, **base)
    b = dns.Simulation(members=mem, **base)
    a.init_fields()
    b.init_fields()
    a.step(10, graph=False)
    b.step(10, graph=True)
    for m in range(len(mem)):
        for f, g in zip(a.get_fields(member=m), b.get_fields(member=m)):
            assert np.array_equal(f, g)


def test_batch_members_match_fp64_oracle(dns):
    """Each member of a batch against its own oracle run (32x32x33 grid, 40 steps)."""
    mem = _members(33)
    batch = dns.Simulation(members=mem, **CHANNEL32)
    batch.init_fields()
    oracles = []
    for m in mem:
        o = orc.Oracle(**{**CHANNEL32, **m})
        o.init_fields()
        oracles.append(o)
    # identical initial fields: the oracle's own initial condition
    for b, o in enumerate(oracles):
        batch.set_fields(*o.get_fields(), member=b)
    batch.step(40)
    for b, o in enumerate(oracles):
        o.step(40)
        e = field_errors(batch.get_fields(member=b), o.get_fields())
        _say(f"member {b} vs oracle after 40 steps: {e}")
        check_parity(e, 1e-10, W_SELF_FP64)
        d, od = batch.diagnostics(member=b), o.diagnostics()
        assert abs(d["pressure_gradient"] - od["pressure_gradient"]) < 1e-12 * max(1.0, abs(od["pressure_gradient"]))


def test_member_index_checked(dns):
    s = dns.Simulation(members=_members(33)[:2], **CHANNEL32)
    with pytest.raises(dns.DnsError):
        s.get_fields(member=2)
    with pytest.raises(dns.DnsError):
        s.diagnostics(member=-1)
    assert s.members == 2 and s.device_bytes > 0
