"""Serving loop (continuous batching) -- scheduling logic on the CPU with a
stand-in backend that counts steps, and the request generator.  The GPU
behaviour (results alone == in the mix) is in test_gpu_serve.py."""
import numpy as np
import pytest

from fortran_loadfile_mitiwing_amd import cases, lbmib
from fortran_loadfile_mitiwing_amd.serve import Server


class FakeCase:
    def __init__(self, name, nx, steps, nbytes):
        self.name = name
        self.params = type("P", (), {"nx": nx, "ny": nx // 2})()
        self.steps = steps
        self.nbytes = nbytes


class FakeSim:
    def __init__(self, case):
        self.case = case
        self.done = 0


class FakeBackend:
    """Steps every member by min(window, remaining), as lbmib_batch_step does."""

    def __init__(self):
        self.members = []
        self.log = []

    def estimate_bytes(self, case):
        return case.nbytes

    def open(self, case):
        return FakeSim(case)

    def device_bytes(self, sim):
        return sim.case.nbytes

    def add(self, sim):
        for m in self.members:
            assert m.done % lbmib.GEOMETRY_BATCH == sim.done % lbmib.GEOMETRY_BATCH, \
                "joins must stay aligned to the geometry batch"
        self.members.append(sim)
        self.log.append(("add", sim.case.name))

    def remove(self, sim):
        self.members.remove(sim)
        self.log.append(("remove", sim.case.name))

    def step(self, n):
        for m in self.members:
            m.done = min(m.case.steps, m.done + n)

    def steps_done(self, sim):
        return sim.done

    def max_steps(self, sim):
        return sim.case.steps

    def result(self, sim):
        return np.zeros((sim.done, 16))

    def close_case(self, sim):
        pass

    def close(self):
        pass


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t


def test_admission_respects_the_memory_budget_and_fifo_order():
    be = FakeBackend()
    clk = Clock()
    s = Server(be, hbm_budget=100, window=16, clock=clk)
    for k, nb in enumerate([60, 30, 50, 10]):
        s.submit(FakeCase("c%d" % k, 512, 40, nb), tag=k)
    s.admit()
    # c0 + c1 fit (90); c2 does not, and FIFO keeps c3 waiting behind it
    assert [j.tag for j in s.running] == [0, 1]
    assert [j.tag for j in s.queue] == [2, 3]
    assert s.resident_bytes() == 90


def test_cases_retire_when_their_steps_are_done_and_free_memory():
    be = FakeBackend()
    clk = Clock()
    s = Server(be, hbm_budget=100, window=16, clock=clk)
    s.submit(FakeCase("long", 512, 64, 60), tag="long")
    s.submit(FakeCase("short", 512, 20, 30), tag="short")
    s.submit(FakeCase("late", 512, 24, 40), tag="late")
    out = []
    for _ in range(3):
        clk.t += 1.0
        out += s.step_window()
    # short finishes in window 2 (20 steps); late is admitted only after that
    assert [r.tag for r in out] == ["short"]
    assert ("add", "late") in be.log and be.log.index(("add", "late")) > be.log.index(("remove", "short"))
    while not s.idle():
        clk.t += 1.0
        out += s.step_window()
    assert sorted(r.tag for r in out) == ["late", "long", "short"]
    assert {r.tag: r.steps for r in out} == {"long": 64, "short": 20, "late": 24}
    assert all(r.finished >= r.admitted >= r.arrival for r in out)


def test_max_cases_and_window_checks():
    be = FakeBackend()
    s = Server(be, hbm_budget=10 ** 9, max_cases=2, window=16, clock=Clock())
    for k in range(5):
        s.submit(FakeCase("c%d" % k, 512, 8, 1), tag=k)
    s.admit()
    assert len(s.running) == 2
    with pytest.raises(ValueError):
        Server(be, hbm_budget=1, window=8)
    with pytest.raises(ValueError):
        s.submit(FakeCase("huge", 512, 8, 10 ** 10))


def test_run_serves_requests_on_the_server_clock():
    be = FakeBackend()
    clk = Clock()

    def sleep(dt):
        clk.t += dt

    class Req:
        def __init__(self, k, t, steps):
            self.case = FakeCase("r%d" % k, 512, steps, 10)
            self.arrival = t
            self.index = k

    s = Server(be, hbm_budget=1000, window=16, clock=clk)
    orig_step = be.step

    def timed_step(n):
        clk.t += 0.01
        orig_step(n)

    be.step = timed_step
    reqs = [Req(0, 0.0, 100), Req(1, 0.05, 30), Req(2, 5.0, 40)]
    res = s.run(reqs, sleep=sleep)
    assert sorted(r.tag for r in res) == [0, 1, 2]
    st = s.stats()
    assert st["requests"] == 3 and st["peak_running_cases"] == 2
    r2 = [r for r in res if r.tag == 2][0]
    assert r2.admitted >= 5.0  # not admitted before it arrived


def test_serving_requests_generator():
    reqs = cases.serving_requests(20, rate_per_s=2.0, nx_max=4096)
    assert len(reqs) == 20
    t = np.array([r.arrival for r in reqs])
    assert np.all(np.diff(t) > 0)
    for r in reqs:
        p = r.case.params
        assert 512 <= p.nx <= 4096 and p.nx % 128 == 0 and p.ny == p.nx // 2
        assert 2000 <= r.case.n_steps <= 20000
        assert r.case.motion.shape[0] == r.case.n_steps
    # same seed, same requests
    again = cases.serving_requests(20, rate_per_s=2.0, nx_max=4096)
    assert [r.case.name for r in again] == [r.case.name for r in reqs]


def test_sweep_cases_follow_the_stated_distribution():
    cs = cases.sweep_cases(24, n_steps=16)
    for c in cs:
        p = c.params
        assert p.nx in range(512, 2049, 128) and p.ny == p.nx // 2
        assert p.pts_per_wing - 1 == p.nx // 4  # 4 markers per lattice unit of chord nx/16
        assert abs(1.0 / p.scale - p.nx / 16.0) < 1e-12
