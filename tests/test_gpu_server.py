"""Serving loop: jobs admitted by HBM budget into slots of a ragged ensemble, advanced
together tick by tick.  A job's result must be the same alone and in the mix: the fields of
every finished job are compared bitwise with the same run alone (dns_create + dns_step)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dns():
    import f90_dns_amd
    f90_dns_amd.lib()
    return f90_dns_amd


def _job(i, nx, ny, nz, precision=1):
    from f90_dns_amd.workloads import BASE
    re = 120.0 + 37.0 * i
    return dict(BASE, nx=nx, ny=ny, nz=nz, re=re, target_pressure_gradient=2.0 / re,
                perturbation_amplitude=0.006 * (1.5 ** (i % 6)), seed=300 + i, precision=precision)


def _alone(dns, p, nsteps):
    s = dns.Simulation(**p)
    s.init_fields()
    s.step(nsteps, graph=True)
    f = s.get_fields()
    s.close()
    return f


# (the nz = 129 job's per-mode solves walk the packed mode space)
JOBS = [((32, 32, 33), 40), ((64, 32, 65), 25), ((32, 32, 33), 60), ((96, 64, 33), 30), ((32, 96, 65), 45),
        ((64, 32, 65), 35), ((32, 24, 129), 20), ((32, 32, 33), 20)]


@pytest.mark.parametrize("precision", [1, 0, 2])
def test_result_alone_equals_result_in_the_mix(dns, precision):
    srv = dns.Server(hbm_budget_bytes=8 << 30, slots_per_group=2)
    ids = []
    # staggered arrivals: later jobs join while earlier ones run
    for i, (shape, n) in enumerate(JOBS):
        ids.append((srv.submit(n, keep_fields=True, **_job(i, *shape, precision=precision)), i, shape, n))
        for _ in range(7):
            srv.tick()
    srv.drain()
    st = srv.stats()
    _say(f"server: {st}")
    assert st["done"] == len(JOBS) and st["failed"] == 0 and st["rejected"] == 0
    assert st["job_steps"] == sum(n for _, n in JOBS)
    assert st["max_running"] >= 3
    assert st["step_p99_ms"] >= st["step_p50_ms"] > 0
    for jid, i, shape, n in ids:
        state, done = srv.status(jid)
        assert state == "done" and done == n
        got, diag = srv.result(jid)
        ref = _alone(dns, _job(i, *shape, precision=precision), n)
        for f in range(4):
            assert np.array_equal(got[f], ref[f]), (jid, f, float(np.max(np.abs(got[f] - ref[f]))))
        assert diag["step"] == n and np.isfinite(diag["energy"])
        t = srv.times(jid)
        assert t["steps"] == n and t["turnaround_s"] >= t["queue_s"] >= 0
    srv.close()


def test_budget_admission_and_rejection(dns):
    one = dns.estimate_bytes(1, **_job(0, 64, 64, 65))
    srv = dns.Server(hbm_budget_bytes=int(one * 1.5), slots_per_group=4)
    big = srv.submit(10, **_job(9, 192, 192, 129))
    ids = [srv.submit(12 + 3 * i, keep_fields=True, **_job(i, 64, 64, 65)) for i in range(3)]
    srv.tick()
    assert srv.status(big)[0] == "rejected"
    assert sum(srv.status(j)[0] == "running" for j in ids) == 1   # one slot fits the budget
    srv.drain()
    st = srv.stats()
    assert st["done"] == 3 and st["rejected"] == 1 and st["peak_bytes"] <= int(one * 1.5)
    for i, jid in enumerate(ids):
        got, _ = srv.result(jid)
        ref = _alone(dns, _job(i, 64, 64, 65), 12 + 3 * i)
        for f in range(4):
            assert np.array_equal(got[f], ref[f])
    srv.close()


def test_idle_group_is_reused_then_released_for_a_new_shape(dns):
    """A group whose jobs have finished stays allocated: a later job of its shape takes its
    slot without a new allocation, and a job of another shape that needs the budget gets it
    once the idle group is released.  Results stay bitwise those of the runs alone."""
    a = dns.estimate_bytes(1, **_job(0, 64, 64, 65))
    b = dns.estimate_bytes(1, **_job(1, 96, 64, 65))
    budget = int(max(a, b) * 1.2)   # one group of one slot at a time
    srv = dns.Server(hbm_budget_bytes=budget, slots_per_group=1)
    first = srv.submit(8, keep_fields=True, **_job(0, 64, 64, 65))
    srv.drain()
    held = srv.stats()["bytes"]
    assert held == a   # idle, still allocated
    again = srv.submit(6, keep_fields=True, **_job(2, 64, 64, 65))
    srv.drain()
    assert srv.stats()["bytes"] == a   # the idle group's slot was reused
    other = srv.submit(5, keep_fields=True, **_job(1, 96, 64, 65))
    srv.drain()
    st = srv.stats()
    assert st["bytes"] == b and st["peak_bytes"] <= budget   # the idle group gave its memory back
    for jid, i, shape, n in ((first, 0, (64, 64, 65), 8), (again, 2, (64, 64, 65), 6), (other, 1, (96, 64, 65), 5)):
        assert srv.status(jid) == ("done", n)
        got, _ = srv.result(jid)
        ref = _alone(dns, _job(i, *shape), n)
        for f in range(4):
            assert np.array_equal(got[f], ref[f]), (jid, f)
    srv.release_idle()
    assert srv.stats()["bytes"] == 0
    srv.close()


def test_non_finite_job_is_evicted_alone(dns):
    """A job whose initial field is non-finite (infinite perturbation amplitude, set on the
    host) is evicted at the next health check; the others finish with their alone results."""
    srv = dns.Server(hbm_budget_bytes=4 << 30, slots_per_group=4)
    srv.set_health_interval(4)
    good = [srv.submit(30, keep_fields=True, **_job(i, 32, 32, 33)) for i in range(3)]
    bad_p = _job(7, 32, 32, 33)
    bad_p["perturbation_amplitude"] = float("inf")
    bad = srv.submit(30, keep_fields=True, **bad_p)
    srv.drain()
    assert srv.status(bad)[0] == "failed", srv.status(bad)
    assert "non-finite" in srv.message(bad)
    for i, jid in enumerate(good):
        assert srv.status(jid) == ("done", 30)
        got, _ = srv.result(jid)
        ref = _alone(dns, _job(i, 32, 32, 33), 30)
        for f in range(4):
            assert np.array_equal(got[f], ref[f])
    srv.close()
