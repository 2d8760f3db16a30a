"""Ragged ensembles: members of different grids, one launch per stage and kernel class.

The bar is the one the batched context meets: each member of an ensemble is bitwise the same
run alone (dns_create of that member), whatever the launch tuning; members are also checked
against the FP64 oracle, and the autotuner must leave every member's state as it found it."""
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from tests.gpu_util import W_SELF_FP64, W_SELF_MIXED, check_parity, field_errors, max_err

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dns():
    import f90_dns_amd
    f90_dns_amd.lib()
    return f90_dns_amd


def _mixed_shapes(precision=0):
    from f90_dns_amd.workloads import BASE
    shapes = [(32, 32, 33), (64, 32, 33), (32, 96, 65), (96, 96, 65), (64, 64, 33), (32, 32, 33), (96, 64, 65)]
    out = []
    for i, (nx, ny, nz) in enumerate(shapes):
        re = 110.0 + 40.0 * i
        out.append(dict(BASE, nx=nx, ny=ny, nz=nz, re=re, target_pressure_gradient=2.0 / re,
                        perturbation_amplitude=0.005 * (1.6 ** i), seed=77 + i, precision=precision,
                        flow_control_method=2 if i % 3 == 1 else 1, target_bulk_velocity=2.0 / 3.0,
                        controller_gain=0.15, controller_update_freq=7))
    return out


def _alone(dns, m, steps):
    s = dns.Simulation(**m)
    s.init_fields()
    s.step(steps, graph=True)
    f = s.get_fields()
    d = s.diagnostics()
    s.close()
    return f, d


def _check_bitwise(dns, ens, members, steps):
    for i, m in enumerate(members):
        ref, dref = _alone(dns, m, steps)
        got = ens.get_fields(i)
        for f in range(4):
            assert np.array_equal(got[f], ref[f]), (i, f, float(np.max(np.abs(got[f] - ref[f]))))
        d = ens.diagnostics(i)
        for k in ("umax", "vmax", "wmax", "energy", "pressure_gradient", "bulk_velocity"):
            assert d[k] == dref[k], (i, k, d[k], dref[k])


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_ensemble_member_equals_member_alone(dns, precision):
    mem = _mixed_shapes(precision)
    ens = dns.Ensemble(mem)
    assert ens.groups == 6 and ens.members == len(mem)
    ens.init_fields()
    ens.step(15, graph=True)
    ens.synchronize()
    _say(f"ensemble of {len(mem)} ({ens.groups} groups), precision {precision}: 15 steps")
    _check_bitwise(dns, ens, mem, 15)
    ens.close()


@pytest.mark.parametrize("precision,waves", [(0, None), (1, None), (2, None), (0, (2, 2, 2)), (1, (4, 4, 4)),
                                             (2, (2, 2, 2))])
def test_even_nz_members_equal_members_alone(dns, precision, waves):
    """Members of even nz (no centre node; one- to four-wave contraction tiles, two row-block
    pairs per wave at nz = 130) beside an odd one: each is bitwise its run alone, also at other
    tile widths."""
    from f90_dns_amd.workloads import BASE
    shapes = [(32, 32, 18), (48, 36, 34), (32, 64, 48), (64, 32, 66), (48, 36, 130), (32, 32, 33), (48, 32, 128)]
    mem = []
    for i, (nx, ny, nz) in enumerate(shapes):
        re = 120.0 + 50.0 * i
        mem.append(dict(BASE, nx=nx, ny=ny, nz=nz, re=re, target_pressure_gradient=2.0 / re, seed=31 + i,
                        perturbation_amplitude=0.01 * (1.5 ** i), precision=precision, lgl_derivative=1,
                        flow_control_method=2 if i % 2 else 1, target_bulk_velocity=2.0 / 3.0, controller_gain=0.15,
                        controller_update_freq=7))
    ens = dns.Ensemble(mem)
    if waves:
        ens.set_tile_waves(*waves)
    ens.init_fields()
    ens.step(10, graph=True)
    ens.synchronize()
    _say(f"even-nz ensemble ({ens.groups} groups), precision {precision}, waves {waves}: 10 steps")
    _check_bitwise(dns, ens, mem, 10)
    ens.close()


@pytest.mark.parametrize("tuning,waves", [(dict(lanes=2), None), (dict(interleave=1), None),
                                          (dict(promote_fp64=(1, 1, 0)), None),
                                          (dict(lanes=2, promote_fp64=(2, 1, 0), interleave=1), None),
                                          (dict(), (2, 2, 2)), (dict(), (4, 4, 4)), (dict(lanes=2), (2, 4, 2)),
                                          (dict(lanes=3), None), (dict(lanes=4), (0, 0, 2))],
                         ids=["lanes2", "interleave", "promote", "all", "waves2", "waves4", "waves_mixed", "lanes3",
                              "lanes4"])
def test_ensemble_tuning_does_not_change_results(dns, tuning, waves):
    """Launch tuning, including the tile shapes of the wall-normal kernels (nz = 33 and 65 groups
    at two or four waves per column block, sharing one launch per stage), leaves every member
    bitwise its run alone."""
    mem = _mixed_shapes(0)
    ens = dns.Ensemble(mem)
    ens.set_tuning(**tuning)
    if waves:
        ens.set_tile_waves(*waves)
    ens.init_fields()
    ens.step(6, graph=True)
    ens.step(4, graph=False)
    _check_bitwise(dns, ens, mem, 10)


def test_one_launch_per_stage(dns):
    """Every plane class of the mix promoted to one class (the 96 x 96 planes' class 5) and the
    wall-normal tiles at two waves: all members share one launch per stage, and each stays
    bitwise its run alone."""
    mem = _mixed_shapes(1)
    ens = dns.Ensemble(mem)
    natural = ens.launch_counts()
    assert natural["c2r"] == natural["r2c"] == 4 and natural["zderiv"] >= 1
    order = dns.Ensemble.SIZE_CLASS_ORDER
    levels = [order.index(5) - order.index(c) if c in (0, 1, 2) else 0 for c in range(7)]
    ens.set_promotion(1, levels)
    ens.set_tile_waves(2, 2, 2)
    counts = ens.launch_counts()
    _say(f"launches per stage: natural {natural}, unified {counts}")
    assert counts == dict(zderiv=1, c2r=1, r2c=1, helmholtz=1, poisson=1)
    ens.init_fields()
    ens.step(6, graph=True)
    ens.step(2, graph=False)
    _check_bitwise(dns, ens, mem, 8)


def test_autotune_restores_state_and_keeps_results(dns):
    mem = _mixed_shapes(1)
    ens = dns.Ensemble(mem)
    ens.init_fields()
    ens.step(3)
    before = [ens.get_fields(i) for i in range(len(mem))]
    rep = ens.autotune(steps_per_trial=3)
    _say(f"autotune: {rep['chosen']} {rep['ms_per_step']:.3f} ms/step, {len(rep['trials'])} trials")
    assert len(rep["trials"]) >= 3
    after = [ens.get_fields(i) for i in range(len(mem))]
    for a, b in zip(before, after):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
    ens.step(7)
    _check_bitwise(dns, ens, mem, 10)


def test_ensemble_members_match_fp64_oracle(dns):
    """The two smallest members against their own oracle runs, from the oracle's initial fields."""
    mem = [m for m in _mixed_shapes(0) if m["nx"] * m["ny"] * m["nz"] <= 64 * 32 * 33]
    ens = dns.Ensemble(mem)
    oracles = []
    for i, m in enumerate(mem):
        o = orc.Oracle(**{k: v for k, v in m.items() if k != "precision"})
        o.init_fields()
        ens.set_fields(i, *o.get_fields())
        oracles.append(o)
    ens.step(30)
    for i, o in enumerate(oracles):
        o.step(30)
        e = field_errors(ens.get_fields(i), o.get_fields())
        _say(f"member {i} {ens.shape(i)} vs oracle after 30 steps: {e}")
        check_parity(e, 1e-10, W_SELF_FP64)


def test_estimate_bytes_matches_allocation(dns):
    for shape in [dict(nx=32, ny=32, nz=33), dict(nx=96, ny=64, nz=65, lgl_derivative=1),
                  dict(nx=128, ny=128, nz=129, lgl_derivative=1, precision=1)]:
        for B in (1, 3):
            est = dns.estimate_bytes(nmembers=B, **shape)
            s = dns.Simulation(members=[{}] * B, **shape)
            assert s.device_bytes == est, (shape, B, s.device_bytes, est)
            s.close()


def test_ensemble_of_64_ragged_members(dns):
    """The 64-member mixed-precision sweep: every member bitwise equal to its run alone."""
    from f90_dns_amd.workloads import ensemble_members
    mem = ensemble_members(64)
    ens = dns.Ensemble(mem)
    _say(f"64-member ensemble: {ens.groups} groups, {ens.device_bytes / 2**30:.2f} GiB")
    ens.init_fields()
    ens.step(8, graph=True)
    _check_bitwise(dns, ens, mem, 8)
    for i in range(len(mem)):
        assert np.isfinite(ens.diagnostics(i)["energy"])


@pytest.mark.parametrize("precision", [0, 1])
def test_ensemble_with_packed_members_equals_alone(dns, precision):
    """Members at nz = 129 (their per-mode solves walk the packed mode space, through the ragged
    kernels' tile map) next to narrow members: each bitwise its run alone."""
    from f90_dns_amd.workloads import BASE
    mem = [dict(BASE, nx=32, ny=24, nz=129, lgl_derivative=1, re=180.0, seed=31, precision=precision),
           dict(BASE, nx=32, ny=32, nz=33, re=150.0, seed=32, precision=precision),
           dict(BASE, nx=48, ny=36, nz=129, lgl_derivative=1, re=240.0, seed=33, flow_control_method=2,
                target_bulk_velocity=2.0 / 3.0, controller_gain=0.15, controller_update_freq=7, precision=precision)]
    ens = dns.Ensemble(mem)
    ens.init_fields()
    ens.step(6, graph=True)
    ens.synchronize()
    _say(f"ensemble with packed members, precision {precision}: 6 steps")
    _check_bitwise(dns, ens, mem, 6)
    ens.close()
