"""The server's admission and retirement logic on the CPU: a stand-in batch
and device, and a simulated clock, so that every path of Server.run (idle
waits, budget and slot limits, rejected jobs, runs queued ahead) is driven
without a GPU."""
from __future__ import annotations

import pytest

pytest.importorskip("torch")

from probabilistic_fluid_simulation_amd import scenes, serve  # noqa: E402


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, d):
        self.t += max(d, 1e-6)  # time always moves on


class Event:
    def __init__(self, clock, at):
        self.clock, self.at = clock, at

    def query(self):
        return self.clock.t >= self.at

    def synchronize(self):
        self.clock.t = max(self.clock.t, self.at)


class FakeBatch:
    """Records which slots every queued step advanced."""

    def __init__(self):
        self.slots = {}
        self.next = 0
        self.taken = {}  # job id -> steps it was advanced
        self.queued = []

    def add(self, vp, tmp, image, itmp, dt, visc, iters, sync=True):
        assert sync is False
        slot, self.next = self.next, self.next + 1
        self.slots[slot] = vp  # the job
        self.taken[vp.job] = 0
        return slot

    def step(self, n):
        assert n >= 1
        for job in self.slots.values():
            self.taken[job.job] += n
        self.queued.append(n)

    def remove(self, slot, hold=True):
        assert hold is False
        job = self.slots.pop(slot)
        return (job,) * 4

    def image(self, slot):
        return self.slots[slot]

    def sync(self):
        pass

    def close(self):
        pass


class FakeDevice:
    """Runs take 1 ms per step, one after another."""

    def __init__(self, clock, batch):
        self.clock, self.batch = clock, batch
        self.busy_until = 0.0
        self.done = 0

    def make(self, job):
        return (job,) * 4

    def mark(self):
        n = self.batch.queued[self.done]
        self.done += 1
        self.busy_until = max(self.busy_until, self.clock.t) + 1e-3 * n
        return Event(self.clock, self.busy_until)

    def copy_out(self, t):
        return ("result", t.job)


def _jobs(n, rate, side=64, steps=None, seed=7):
    specs = scenes.served_specs(n, rate, seed=seed)
    for k, sp in enumerate(specs):
        sp.update(fw=side, fh=side, iw=side, ih=side)
        if steps is not None:
            sp.update(steps=steps[k % len(steps)])
    return serve.jobs_from_specs(specs)


def _server(clock, **kw):
    b = FakeBatch()
    return serve.Server(batch=b, device=FakeDevice(clock, b), **kw), b


def _run(srv, jobs, clock):
    return srv.run(jobs, clock=clock, sleep=clock.sleep)


def test_every_job_gets_exactly_its_steps():
    clock = Clock()
    srv, b = _server(clock, max_substeps=7, keep_results=True)
    jobs = _jobs(40, rate=200.0, steps=[50, 13, 500, 1, 77])
    out = _run(srv, jobs, clock)
    assert out["completed"] == 40 and out["rejected"] == 0
    for j in jobs:
        assert b.taken[j.job] == j.steps
        assert j.result == ("result", j.job)
        assert j.arrival <= j.admitted <= j.finished
    assert not srv.active and srv.used == 0
    assert all(n <= 7 for n in b.queued)


def test_oversized_job_is_rejected_not_waited_for():
    """A job larger than the whole budget can never be admitted: it is
    rejected and the jobs behind it still run (no livelock)."""
    clock = Clock()
    jobs = _jobs(5, rate=100.0, steps=[20])
    big = jobs[2]
    big.fw = big.fh = big.iw = big.ih = 1024
    srv, b = _server(clock, hbm_budget=jobs[0].nbytes * 3)
    out = _run(srv, jobs, clock)
    assert big.rejected and big.finished is None and big.job not in b.taken
    assert out["rejected"] == 1 and out["completed"] == 4


def test_budget_and_slot_limits():
    clock = Clock()
    jobs = _jobs(12, rate=1e6, steps=[30, 10])
    srv, _ = _server(clock, hbm_budget=2 * jobs[0].nbytes)
    out = _run(srv, jobs, clock)
    assert out["completed"] == 12 and srv.peak_active <= 2
    clock = Clock()
    jobs = _jobs(12, rate=1e6, steps=[30, 10])
    srv, _ = _server(clock, max_active=3)
    out = _run(srv, jobs, clock)
    assert out["completed"] == 12 and srv.peak_active == 3


def test_runs_queued_ahead_are_bounded_by_depth():
    """Runs are queued ahead of the device (more than one unfinished at a
    time) but never more than `depth` of them."""
    clock = Clock()
    jobs = _jobs(6, rate=1e6, steps=[100])
    srv, b = _server(clock, max_substeps=4, depth=3)
    events, ahead = [], []
    orig_mark = srv.dev.mark

    def mark():
        ahead.append(sum(1 for e in events if not e.query()))
        e = orig_mark()
        events.append(e)
        return e

    srv.dev.mark = mark
    out = _run(srv, jobs, clock)
    assert out["completed"] == 6
    assert max(ahead) == 2 and all(a < 3 for a in ahead)
    assert sum(b.queued) >= 100 and all(n <= 4 for n in b.queued)


def test_idle_server_waits_for_the_next_arrival():
    clock = Clock()
    jobs = _jobs(3, rate=0.5, steps=[5])  # arrivals seconds apart
    srv, _ = _server(clock)
    out = _run(srv, jobs, clock)
    assert out["completed"] == 3
    for j in jobs:
        assert j.admitted >= j.arrival and j.admitted - j.arrival < 1e-3


def test_invalid_limits_raise():
    for kw in ({"max_active": 0}, {"hbm_budget": 0}, {"max_substeps": 0}, {"depth": 0}):
        with pytest.raises(ValueError):
            serve.Server(batch=FakeBatch(), device=object(), **kw)
