"""Serving loop logic on the CPU (a stand-in context replaces the GPU one) and, on the GPU,
that every job's result is the same alone and in the mix."""
import time

import numpy as np
import pytest

import fluam_amd as F
import fluam_amd.serve as S
from fluam_amd._abi import make_params, SCHEME_STOKES_LIMIT, SCHEME_QUASI_NEUTRALLY_BUOYANT


class FakeSim:
    """Completes each step `dt` seconds after it was issued (per context, in order)."""
    live = 0
    peak = 0
    dt = 0.001

    def __init__(self, params):
        self.p = params
        self.t_ready = time.perf_counter()
        self.marker = None
        self.steps = 0
        FakeSim.live += 1
        FakeSim.peak = max(FakeSim.peak, FakeSim.live)

    def set_graph(self, on):
        pass

    def set_noise_overlap(self, on):
        self.overlap = on

    def step(self, n=1):
        self.t_ready = max(self.t_ready, time.perf_counter()) + n * FakeSim.dt
        self.steps += n

    def record(self):
        self.marker = self.t_ready

    def done(self):
        return self.marker is None or time.perf_counter() >= self.marker

    def sync(self):
        while time.perf_counter() < self.t_ready:
            time.sleep(0.0001)

    def velocity(self):
        return np.full(3, float(self.steps))

    def particles(self):
        return np.zeros(3)

    def close(self):
        FakeSim.live -= 1


@pytest.fixture
def fake(monkeypatch):
    FakeSim.live = FakeSim.peak = 0
    monkeypatch.setattr(S, "Simulation", FakeSim)
    monkeypatch.setattr(S, "bytes_needed", lambda p: p.mx * p.my * p.mz)
    return FakeSim


def _job(name, edge, steps, arrival):
    p = make_params(SCHEME_STOKES_LIMIT, cells=(edge, edge, edge))
    return S.Job(name=name, params=p, steps=steps, arrival=arrival)


def test_all_jobs_complete_with_latencies(fake):
    jobs = [_job(f"j{i}", 16, 20 + 5 * i, 0.002 * i) for i in range(6)]
    r = S.Server(budget=10 ** 9, keep_results=True).run(jobs)
    assert r["jobs"] == 6
    for j in jobs:
        assert j.finished is not None and j.finished >= j.admitted >= j.arrival - 1e-9
        assert j.velocity[0] == j.steps  # every step issued exactly once
        # a job cannot finish before its own work: steps x dt after admission
        assert j.finished - j.admitted >= j.steps * FakeSim.dt * 0.99
    assert r["p99_latency_s"] >= r["p50_latency_s"] > 0
    assert fake.live == 0


def test_budget_limits_admission(fake):
    # each job holds 16^3 = 4096 "bytes"; a budget of 10000 admits two at a time
    jobs = [_job(f"j{i}", 16, 30, 0.0) for i in range(5)]
    S.Server(budget=10000).run(jobs)
    assert fake.peak == 2
    order = sorted(jobs, key=lambda j: j.admitted)
    # the third job is admitted only after one of the first two has finished
    first_done = min(j.finished for j in order[:2])
    assert order[2].admitted >= first_done - 1e-6


@pytest.mark.parametrize("merged", [False, True])
def test_max_active_caps_running_jobs(fake, monkeypatch, merged):
    # the memory budget would hold all six; the cap admits three at a time, the rest as they leave
    if merged:
        monkeypatch.setattr(S, "Group", FakeGroup)
    jobs = [_job(f"j{i}", 16, 20 + 3 * i, 0.0) for i in range(6)]
    r = S.Server(budget=10 ** 9, max_active=3, merged=merged).run(jobs)
    assert r["jobs"] == 6
    assert fake.peak == 3
    order = sorted(jobs, key=lambda j: j.admitted)
    assert order[3].admitted >= min(j.finished for j in order[:3]) - 1e-6


def test_job_larger_than_budget_is_rejected(fake):
    with pytest.raises(ValueError):
        S.Server(budget=100).run([_job("big", 16, 5, 0.0)])


def test_jobs_overlap_continuous_batching(fake):
    # a job arriving while another runs is admitted at the next step boundary, not after it
    jobs = [_job("long", 16, 200, 0.0), _job("short", 16, 10, 0.01)]
    S.Server(budget=10 ** 9).run(jobs)
    long_, short = jobs
    assert short.admitted < long_.finished
    assert short.finished < long_.finished


def test_device_error_evicts_only_that_job(fake, monkeypatch):
    """A job whose steps raise a device error flag (fhd_query reports -1, raised as FhdError by
    done()) is evicted with the reason recorded; the other jobs run to completion."""
    class Failing(FakeSim):
        def done(self):
            if self.p.mx == 24 and self.steps >= 8:
                raise F.FhdError("device error: non-finite particle displacement")
            return super().done()

    monkeypatch.setattr(S, "Simulation", Failing)
    jobs = [_job("a", 16, 40, 0.0), _job("bad", 24, 40, 0.0), _job("c", 16, 30, 0.001)]
    r = S.Server(budget=10 ** 9, keep_results=True, max_inflight=2).run(jobs)
    a, bad, c = jobs
    assert bad.error and "non-finite" in bad.error
    assert bad.velocity is None and bad.finished is not None
    assert a.error is None and c.error is None
    assert a.velocity[0] == 40 and c.velocity[0] == 30
    assert r["jobs"] == 2 and r["evicted"] == 1
    assert Failing.live == 0


class FakeGroup:
    """Stand-in merged-stage batch over FakeSims: a group step advances every member."""
    built = 0

    def __init__(self, sims):
        self.replicas = list(sims)
        self.open = True
        FakeGroup.built += 1

    def set_mode(self, m):
        assert m == 2

    def step(self, n=1):
        assert self.open
        for s in self.replicas:
            s.step(n)

    def sync(self):
        err = None
        for s in self.replicas:
            try:
                s.sync()
            except F.FhdError as e:
                err = e
        if err:
            raise err

    def close(self):
        self.open = False


@pytest.fixture
def fake_group(fake, monkeypatch):
    FakeGroup.built = 0
    monkeypatch.setattr(S, "Group", FakeGroup)
    return FakeGroup


def test_merged_jobs_join_and_leave(fake_group):
    """Jobs of different lengths arriving at different times each take exactly their own step
    count; the group is rebuilt as they join and leave (also when nothing new arrives)."""
    jobs = [_job("a", 16, 50, 0.0), _job("b", 16, 9, 0.0), _job("c", 20, 23, 0.005), _job("d", 16, 4, 0.2)]
    r = S.Server(budget=10 ** 9, keep_results=True, merged=True, chunk=8).run(jobs)
    assert r["jobs"] == 4 and r["evicted"] == 0
    for j in jobs:
        assert j.velocity[0] == j.steps, j.name
        assert j.finished is not None and j.finished >= j.admitted
    assert fake_group.built >= 4
    assert FakeSim.live == 0


def test_merged_qnb_jobs_bootstrap_alone(fake_group):
    p = make_params(SCHEME_QUASI_NEUTRALLY_BUOYANT, cells=(16, 16, 16), np=10)
    jobs = [S.Job("q", p, 11, 0.0), _job("s", 16, 6, 0.0), S.Job("q1", p, 1, 0.0)]
    S.Server(budget=10 ** 9, keep_results=True, merged=True).run(jobs)
    assert [j.velocity[0] for j in jobs] == [11, 6, 1]


def test_merged_device_error_evicts_only_that_job(fake_group, monkeypatch):
    class Failing(FakeSim):
        def sync(self):
            super().sync()
            if self.p.mx == 24 and self.steps >= 8:
                raise F.FhdError("device error: non-finite particle displacement")

    monkeypatch.setattr(S, "Simulation", Failing)
    jobs = [_job("a", 16, 40, 0.0), _job("bad", 24, 40, 0.0), _job("c", 16, 30, 0.001)]
    r = S.Server(budget=10 ** 9, keep_results=True, merged=True, chunk=4).run(jobs)
    a, bad, c = jobs
    assert bad.error and "non-finite" in bad.error and bad.velocity is None
    assert a.velocity[0] == 40 and c.velocity[0] == 30
    assert r["jobs"] == 2 and r["evicted"] == 1
    assert Failing.live == 0


def test_poisson_arrivals_are_increasing():
    jobs = S.poisson_jobs(50.0, 20, 3, lambda i, rng: (make_params(SCHEME_STOKES_LIMIT, cells=(16, 16, 16)), 5))
    t = [j.arrival for j in jobs]
    assert all(b > a for a, b in zip(t, t[1:]))


@pytest.mark.gpu
@pytest.mark.parametrize("merged", [False, True])
def test_served_results_equal_alone(merged):
    """Jobs of different shapes and schemes served together give the same bits as each alone,
    with per-job streams and with the running jobs regrouped into one merged-stage batch."""
    def params(i, edge, np_, scheme):
        return make_params(scheme, cells=(edge, edge, edge), np=np_, kernel_points=3, temperature=1e-3,
                           dt=0.05, seed=500 + i, initial_positions=1, cutoff=2 ** (1 / 6), wca_sigma=1.0)
    specs = [(32, 200, SCHEME_STOKES_LIMIT, 30), (16, 0, SCHEME_STOKES_LIMIT, 50),
             (24, 100, SCHEME_QUASI_NEUTRALLY_BUOYANT, 40), (32, 0, SCHEME_QUASI_NEUTRALLY_BUOYANT, 20)]
    mix = [S.Job(f"j{i}", params(i, e, n, sc), steps, arrival=0.003 * i) for i, (e, n, sc, steps) in enumerate(specs)]
    res = S.Server(keep_results=True, merged=merged, chunk=7).run(mix)
    assert res["jobs"] == len(specs) and res["evicted"] == 0
    for i, (e, n, sc, steps) in enumerate(specs):
        alone = S.Job(f"a{i}", params(i, e, n, sc), steps)
        S.Server(keep_results=True).run([alone])
        assert np.array_equal(alone.velocity, mix[i].velocity)
        if n:
            assert np.array_equal(alone.particles, mix[i].particles)
