"""Serving pool on the GPU: systems admitted into and retired from fixed
slots while others run, each finishing bitwise where it finishes alone
(same positions and velocities after the same number of steps), slot reuse
by a smaller system, capacity errors, and the scheduler end to end."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def jobs(pm):
    out = []
    for k, (n, seed) in enumerate([(2400, 31), (4100, 32), (1300, 33)]):
        top = pm.Topology.generate(n, seed=seed)
        p = pm.default_params(seed=500 + k, stream_id=k, temperature_K=285.0 + 20 * k,
                              end_temperature_K=285.0 + 20 * k)
        out.append((top, p))
    return out


def alone(pm, top, p, steps):
    with pm.FusedEngine(top, p) as eng:
        eng.step(steps)
        return eng.get_positions(), eng.get_velocities(), eng.stats()["status"]


@pytest.fixture(scope="module")
def pool(pm, jobs):
    template = jobs[1][0]
    caps = [pm.slot_caps_for(template, atoms=6000, scale=1.2) for _ in range(3)]
    with pm.ServingPool(caps) as p:
        yield p


def test_admit_retire_reuse_bitwise_alone(pm, jobs, pool):
    (ta, pa), (tb, pb), (tc, pc) = jobs
    pool.admit(0, ta, pa)
    pool.step(30, 10)
    pool.admit(1, tb, pb)
    pool.step(40, 10)
    xa, va = pool.get_positions(0), pool.get_velocities(0)  # A has run 70 steps
    pool.retire(0)
    pool.admit(0, tc, pc)  # a smaller system reuses A's slot
    pool.step(50, 10)
    xb, vb = pool.get_positions(1), pool.get_velocities(1)  # B: 90 steps
    xc, vc = pool.get_positions(0), pool.get_velocities(0)  # C: 50 steps
    assert pool.stats(0)["steps_done"] == 50 and pool.stats(1)["steps_done"] == 90
    for (x, v), (top, p, n) in zip([(xa, va), (xb, vb), (xc, vc)], [(ta, pa, 70), (tb, pb, 90), (tc, pc, 50)]):
        x1, v1, status = alone(pm, top, p, n)
        assert status == 0
        assert np.array_equal(x, x1) and np.array_equal(v, v1)
    pool.retire(0)
    pool.retire(1)


def test_capacity_and_occupancy_errors(pm, jobs, pool):
    big = pm.Topology.generate(7000, seed=40)
    with pytest.raises(pm.PrismMdError) as ex:
        pool.admit(2, big, pm.default_params())
    assert ex.value.rc == pm.E_CAPACITY
    assert not pool.occupied(2)
    pool.admit(2, jobs[2][0], jobs[2][1])
    with pytest.raises(pm.PrismMdError) as ex:
        pool.admit(2, jobs[2][0], jobs[2][1])
    assert ex.value.rc == pm.E_INVALID
    pool.retire(2)
    assert not pool.occupied(2)


def test_scheduler_end_to_end(pm, jobs, pool):
    from prism_md.serving import Job, Scheduler
    sch = Scheduler(pool, graph_steps=10)
    lengths = [57, 100, 43, 80]  # two end inside a replay
    for k, steps in enumerate(lengths):
        top, p = jobs[k % 3]
        sch.submit(Job(k, top, p, steps))
    while sch.tick():
        pass
    rep = sch.report(1.0)
    assert len(sch.done) == 4 and not rep["faulted_jobs"]
    assert rep["step_latency_ms"]["samples"] == sum(n - 1 for n in lengths)
    for job, n in zip(sorted(sch.done, key=lambda j: j.jid), lengths):
        assert job.done_steps == n
        x1, v1, _ = alone(pm, job.topology, job.params, n)
        assert np.array_equal(job.final_positions, x1), job.jid


def test_step_limit_stops_inside_a_replay(pm, jobs, pool):
    """A step limit of 17 inside 30 steps of 10-step graph replays: the job
    stops at step 17 exactly, bitwise where a 17-step run alone ends, while
    the job beside it runs all 30; a single engine stops alike."""
    (ta, pa), (tb, pb) = jobs[0], jobs[1]
    pool.admit(0, ta, pa)
    pool.admit(1, tb, pb)
    pool.set_step_limit(0, 17)
    pool.step(30, 10)
    assert pool.stats(0)["steps_done"] == 17 and pool.stats(1)["steps_done"] == 30
    xa, va, _ = alone(pm, ta, pa, 17)
    assert np.array_equal(pool.get_positions(0), xa) and np.array_equal(pool.get_velocities(0), va)
    pool.retire(0)
    pool.retire(1)
    with pm.FusedEngine(ta, pa) as eng:
        eng.set_step_limit(25)
        eng.step(40)
        assert eng.stats()["steps_done"] == 25
        eng.set_step_limit(0)
        eng.step(5)
        assert eng.stats()["steps_done"] == 30


def test_device_step_latency(pm, jobs, pool):
    """Per-step latency from the device's step-end stamps: one sample per
    step after a job's first, each the time since the previous step's end;
    an admission between two ticks shows up as a long interval of the job
    that kept running (the admission's device and host work sit between its
    two steps)."""
    import time
    (ta, pa), (tb, pb) = jobs[0], jobs[1]
    pool.admit(0, ta, pa)
    t0 = time.perf_counter()
    pool.step(40, 10)
    host_ms = (time.perf_counter() - t0) * 1e3 / 40
    lat = pool.latency_samples(0)
    print(f"[latency] {len(lat)} samples, p50 {np.median(lat):.4f} ms, host {host_ms:.4f} ms/step", flush=True)
    assert len(lat) == 39 and (lat > 0).all()
    assert np.median(lat) <= 2.0 * host_ms
    pool.admit(1, tb, pb)  # lowering, upload, re-sort and forces of the new job, between two ticks
    pool.step(10, 10)
    lat = pool.latency_samples(0)
    gap = lat[39]  # the first step after the admission
    print(f"[latency] step after the admission {gap:.4f} ms vs median {np.median(lat[:39]):.4f} ms", flush=True)
    assert len(lat) == 49 and gap > 5.0 * np.median(lat[:39])
    q = pool.latency(0)
    assert q["samples"] == 49 and q["p99_ms"] >= q["p50_ms"] > 0
    assert len(pool.latency_samples(1)) == 9
    pool.retire(0)
    pool.retire(1)


def test_faulted_job_halts_alone(pm, jobs, pool):
    """A job whose SHAKE cannot converge (one Newton sweep against a 1e-14
    tolerance) halts at the end of its first step and is retired by the
    scheduler; the job beside it runs to its length bitwise as it runs alone."""
    from prism_md.serving import Job, Scheduler
    top_bad = pm.Topology.generate(1500, seed=34)
    p_bad = pm.default_params(constraint_max_iter=1, shake_tol=1e-14, seed=77, stream_id=9)
    (tb, pb) = jobs[0]
    sch = Scheduler(pool, graph_steps=10)
    sch.submit(Job(0, top_bad, p_bad, 50))
    sch.submit(Job(1, tb, pb, 60))
    while sch.tick():
        pass
    assert [j.jid for j in sch.failed] == [0] and [j.jid for j in sch.done] == [1]
    bad = sch.failed[0]
    assert bad.done_steps == 1 and bad.status & 1, (bad.done_steps, bad.status, bad.error)
    assert "did not converge" in bad.error
    good = sch.done[0]
    x1, v1, status = alone(pm, tb, pb, 60)
    assert status == 0 and good.status == 0
    assert np.array_equal(good.final_positions, x1) and np.array_equal(good.final_velocities, v1)
    for s in range(3):
        assert not pool.occupied(s)
