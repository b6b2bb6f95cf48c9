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
    lengths = [60, 100, 40, 80]
    for k, steps in enumerate(lengths):
        top, p = jobs[k % 3]
        sch.submit(Job(k, top, p, steps))
    while sch.tick():
        pass
    assert len(sch.done) == 4 and not sch.report(1.0)["faulted_jobs"]
    for job in sch.done:
        x1, v1, _ = alone(pm, job.topology, job.params, job.done_steps)
        assert np.array_equal(job.final_positions, x1), job.jid


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
