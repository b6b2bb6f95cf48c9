"""Host logic of the serving loop (no GPU): slot planning within an HBM
budget, FIFO admission into the smallest fitting free slot, retirement and
slot reuse, latency and throughput accounting -- driven through a stand-in
pool that records the calls."""
from __future__ import annotations

import types

import numpy as np
import pytest

from prism_md.serving import Job, Scheduler, plan_slots


class FakePool:
    """Slots that count steps and stop at their step limit, as the device
    does; each step's latency is 0.25 ms."""

    def __init__(self, atom_caps):
        self.caps = [types.SimpleNamespace(atoms=a) for a in atom_caps]
        self.log = []
        self.occupied = {}
        self.steps = {}
        self.limit = {}

    def fits(self, slot, top):
        return top.n_atoms < self.caps[slot].atoms

    def admit(self, slot, top, params):
        assert slot not in self.occupied
        self.occupied[slot] = top
        self.steps[slot] = 0
        self.limit[slot] = 0
        self.log.append(("admit", slot, top.n_atoms))

    def set_step_limit(self, slot, end_step):
        self.limit[slot] = end_step

    def retire(self, slot):
        del self.occupied[slot]
        self.log.append(("retire", slot, self.steps[slot]))

    def step(self, n, graph_steps):
        self.log.append(("step", n, tuple(sorted(self.occupied))))
        for slot in self.occupied:
            cap = self.limit[slot] or float("inf")
            self.steps[slot] = int(min(cap, self.steps[slot] + n))

    def latency_samples(self, slot):
        return np.full(max(0, self.steps[slot] - 1), 0.25, np.float32)

    def get_positions(self, slot):
        return None

    def get_velocities(self, slot):
        return None

    def stats(self, slot):
        return {"status": 0}


def top(n):
    return types.SimpleNamespace(n_atoms=n)


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 0.001
        return self.t


def test_plan_slots_respects_budget_and_mix():
    counts = plan_slots(1e6, 10.0, [1000, 4000, 16000], [0.5, 0.3, 0.2])
    used = sum(c * 10.0 * n for c, n in zip([1000, 4000, 16000], counts))
    assert used <= 1e6 and all(n >= 1 for n in counts)
    assert counts[0] >= counts[2]  # the mix favours the small class
    with pytest.raises(ValueError):
        plan_slots(1e4, 10.0, [1000, 4000], [0.5, 0.5])


def test_smallest_fitting_slot_fifo_and_reuse():
    pool = FakePool([1000, 5000, 5000])
    sch = Scheduler(pool, graph_steps=10, clock=Clock())
    big, small, mid, tiny = Job(0, top(4000), None, 20), Job(1, top(500), None, 10), Job(2, top(3000), None, 30), \
        Job(3, top(900), None, 10)
    for j in (big, small, mid, tiny):
        sch.submit(j)
    sch.tick()
    # big -> a 5000 slot, small -> the 1000 slot, mid -> the other 5000 slot; tiny waits
    assert big.slot in (1, 2) and small.slot == 0 and mid.slot in (1, 2) and tiny.slot is None
    assert small in sch.done and 0 not in sch.running  # 10 steps: done after one tick
    sch.tick()  # tiny takes the freed 1000 slot
    assert tiny.slot == 0 and big in sch.done
    while sch.tick():
        pass
    assert {j.jid for j in sch.done} == {0, 1, 2, 3}
    # the pool never held two jobs in one slot and every admit was matched by a retire
    admits = [e for e in pool.log if e[0] == "admit"]
    retires = [e for e in pool.log if e[0] == "retire"]
    assert len(admits) == len(retires) == 4
    rep = sch.report(wall=1.0)
    assert rep["jobs_done"] == 4 and rep["step_latency_ms"]["p50"] == pytest.approx(0.25)
    assert rep["step_latency_ms"]["samples"] == sum(j.steps - 1 for j in sch.done)
    assert rep["atom_steps_per_s"] == pytest.approx(sum(j.n_atoms * j.done_steps for j in sch.done))


def test_jobs_end_at_their_exact_length():
    """A job whose length is not a multiple of the replay length stops inside
    a replay (device step limit) and is counted for exactly its steps."""
    pool = FakePool([5000, 5000])
    sch = Scheduler(pool, graph_steps=10, clock=Clock())
    a, b = Job(0, top(1000), None, 25), Job(1, top(2000), None, 40)
    sch.submit(a)
    sch.submit(b)
    while sch.tick():
        pass
    retired = {e[1]: e[2] for e in pool.log if e[0] == "retire"}
    assert retired[a.slot] == 25 and retired[b.slot] == 40
    assert a.done_steps == 25 and b.done_steps == 40
    assert sch.atom_steps == 1000 * 25 + 2000 * 40


def test_oversized_job_waits_without_blocking_others():
    pool = FakePool([1000])
    sch = Scheduler(pool, graph_steps=10, clock=Clock())
    huge, small = Job(0, top(5000), None, 10), Job(1, top(200), None, 10)
    sch.submit(huge)
    sch.submit(small)
    sch.tick()
    assert small.slot == 0 and huge.slot is None and list(sch.queue) == [huge]


class FaultyPool(FakePool):
    """A pool whose step latches a fault in chosen slots at a chosen step."""

    def __init__(self, atom_caps, fault_at):
        super().__init__(atom_caps)
        self.fault_at = fault_at  # slot -> step count at which it halts
        self.steps = {}
        self.status = {}

    def admit(self, slot, top, params):
        super().admit(slot, top, params)
        self.status[slot] = 0

    def step(self, n, graph_steps):
        self.log.append(("step", n, tuple(sorted(self.occupied))))
        err = None
        for slot in self.occupied:
            if self.status[slot]:
                continue
            at = self.fault_at.get(slot)
            cap = self.limit[slot] or float("inf")
            if at is not None and self.steps[slot] + n >= at:
                self.steps[slot] = at
                self.status[slot] = 1
                err = err or RuntimeError(f"constraint iteration did not converge (system {slot})")
            else:
                self.steps[slot] = int(min(cap, self.steps[slot] + n))
        if err is not None:
            raise err

    def stats(self, slot):
        return {"status": self.status[slot], "steps_done": self.steps[slot]}


def test_faulted_job_is_retired_and_others_continue():
    pool = FaultyPool([5000, 5000], fault_at={0: 25})
    sch = Scheduler(pool, graph_steps=10, clock=Clock())
    bad, good, late = Job(0, top(1000), None, 100), Job(1, top(2000), None, 60), Job(2, top(1500), None, 20)
    for j in (bad, good, late):
        sch.submit(j)
    while sch.tick():
        pass
    assert [j.jid for j in sch.failed] == [0] and bad.done_steps == 25 and bad.status == 1
    assert "did not converge" in bad.error
    assert {j.jid for j in sch.done} == {1, 2}  # the freed slot took the late job
    assert late.slot == 0 and good.done_steps == 60
    rep = sch.report(wall=1.0)
    assert rep["faulted_jobs"] == [0] and rep["faults"][0]["at_step"] == 25
    assert rep["atom_steps_per_s"] == pytest.approx(1000 * 25 + 2000 * 60 + 1500 * 20)


def test_failure_naming_no_job_is_raised():
    class Broken(FakePool):
        def step(self, n, graph_steps):
            raise RuntimeError("device lost")

    sch = Scheduler(Broken([1000]), graph_steps=10, clock=Clock())
    sch.submit(Job(0, top(100), None, 10))
    with pytest.raises(RuntimeError, match="device lost"):
        sch.tick()
