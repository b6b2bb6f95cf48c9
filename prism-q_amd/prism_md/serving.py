"""Continuous-batching serving loop over a ServingPool.

Requests (jobs) arrive while others run.  Each is admitted into the smallest
free slot whose capacities fit it, as soon as one is free, with a device step
limit at its length; every tick advances all running jobs by one replay of
the captured step graph (graph_steps steps) -- a job that reaches its length
inside a replay stops there -- and a job that has completed its length is
retired and its slot reused.  Coordinates, velocities, lists and noise
counters stay on the device from admission to retirement.

Per-step latency is the device's: the time between consecutive step ends on
the device clock (prism_md_batch_latency_samples), kept per job from its
admission, so the gaps that admissions and retirements of other jobs put
between two of a job's steps are in it.

The pool is sized up front from an HBM budget (plan_slots): slot classes by
atom capacity, counts following the expected job-size mix, until the budget
is spent.  Reference analogues: the stream-ordered VRAM pool and batch
scheduler (crates/prism-nhs/src/vram_pool.rs:1-40,
crates/prism-nhs/src/batch_scheduler.rs:1-16).
"""
from __future__ import annotations

import collections
import dataclasses
import time
from typing import Callable

import numpy as np


@dataclasses.dataclass
class Job:
    jid: int
    topology: object
    params: object
    steps: int
    t_arrive: float = 0.0
    t_admit: float | None = None
    t_done: float | None = None
    slot: int | None = None
    done_steps: int = 0
    final_positions: np.ndarray | None = None
    final_velocities: np.ndarray | None = None
    status: int = 0
    error: str | None = None
    latency_ms: np.ndarray | None = None   # device per-step latencies of the job

    @property
    def n_atoms(self) -> int:
        return self.topology.n_atoms


def plan_slots(budget_bytes: float, bytes_per_atom: float, classes, weights, min_per_class: int = 1):
    """Slot counts per atom-capacity class within an HBM budget.  Each class
    gets slots in proportion to `weights` (the expected share of jobs that
    need that class) until the next slot would exceed the budget."""
    classes = list(classes)
    w = np.asarray(weights, float)
    w = w / w.sum()
    counts = [min_per_class] * len(classes)
    used = sum(c * bytes_per_atom * n for c, n in zip(classes, counts))
    if used > budget_bytes:
        raise ValueError("budget below one slot per class")
    while True:
        # the class furthest below its share of slots gets the next one
        tot = sum(counts)
        k = int(np.argmax(w * (tot + 1) - np.asarray(counts)))
        cost = classes[k] * bytes_per_atom
        if used + cost > budget_bytes:
            break
        counts[k] += 1
        used += cost
    return counts


class Scheduler:
    """FIFO continuous batching over a pool exposing admit/retire/step/fits."""

    def __init__(self, pool, graph_steps: int = 10, clock: Callable[[], float] = time.perf_counter):
        self.pool = pool
        self.graph_steps = graph_steps
        self.clock = clock
        self.queue: collections.deque[Job] = collections.deque()
        self.running: dict[int, Job] = {}
        self.done: list[Job] = []
        self.failed: list[Job] = []
        self.tick_latency: list[float] = []   # host seconds per step, one entry per tick
        self.atom_steps = 0
        self.busy = 0.0
        self.n_slots = len(pool.caps)

    def submit(self, job: Job):
        job.t_arrive = self.clock()
        self.queue.append(job)

    def _free_slots(self):
        return [s for s in range(self.n_slots) if s not in self.running]

    def admit_ready(self):
        """Admit queued jobs in arrival order; a job that fits no free slot
        waits, and later (smaller) jobs may pass it."""
        free = sorted(self._free_slots(), key=lambda s: self.pool.caps[s].atoms)
        waiting = collections.deque()
        while self.queue:
            job = self.queue.popleft()
            slot = next((s for s in free if self.pool.fits(s, job.topology)), None)
            if slot is None:
                waiting.append(job)
                continue
            free.remove(slot)
            self.pool.admit(slot, job.topology, job.params)
            self.pool.set_step_limit(slot, job.steps)  # stops at its length, even inside a replay
            job.slot, job.t_admit = slot, self.clock()
            self.running[slot] = job
        self.queue = waiting

    def tick(self) -> bool:
        """Admit, advance every running job by graph_steps, retire finished
        jobs.  Returns False when nothing ran."""
        self.admit_ready()
        if not self.running:
            return False
        t0 = self.clock()
        fault = None
        try:
            self.pool.step(self.graph_steps, self.graph_steps)
        except RuntimeError as ex:  # PrismMdError: a latched fault halts that job only
            fault = ex
        dt = self.clock() - t0
        self.busy += dt
        self.tick_latency.append(dt / self.graph_steps)
        if fault is not None:
            self._retire_faulted(fault)
        for slot, job in list(self.running.items()):
            ran = min(self.graph_steps, job.steps - job.done_steps)  # the device limit stops it there
            job.done_steps += ran
            self.atom_steps += ran * job.n_atoms
            if job.done_steps >= job.steps:
                job.final_positions = self.pool.get_positions(slot)
                job.final_velocities = self.pool.get_velocities(slot)
                job.status = self.pool.stats(slot)["status"]
                job.latency_ms = self.pool.latency_samples(slot)
                self.pool.retire(slot)
                job.t_done = self.clock()
                self.done.append(job)
                del self.running[slot]
        return True

    def _retire_faulted(self, fault: Exception):
        """Retire every running job whose system latched a fault this tick
        (it halted at the end of its faulting step; the others ran the whole
        tick).  A failure that names no running job is re-raised."""
        faulted = []
        for slot, job in self.running.items():
            st = self.pool.stats(slot)
            if st["status"] != 0:
                faulted.append((slot, job, st))
        if not faulted:
            raise fault
        for slot, job, st in faulted:
            done_here = max(0, int(st["steps_done"]) - job.done_steps)
            self.atom_steps += done_here * job.n_atoms
            job.done_steps = int(st["steps_done"])
            job.status = int(st["status"])
            job.error = str(fault) if len(faulted) == 1 else f"device status {job.status}"
            job.final_positions = self.pool.get_positions(slot)
            job.final_velocities = self.pool.get_velocities(slot)
            job.latency_ms = self.pool.latency_samples(slot)
            self.pool.retire(slot)
            job.t_done = self.clock()
            self.failed.append(job)
            del self.running[slot]

    def job_latencies(self) -> np.ndarray:
        """Device per-step latencies of every job, finished or running (one
        entry per job-step: a step shared by k jobs counts k times)."""
        parts = [j.latency_ms for j in self.done + self.failed if j.latency_ms is not None]
        parts += [self.pool.latency_samples(slot) for slot in self.running]
        return np.concatenate(parts) if parts else np.zeros(0, np.float32)

    def report(self, wall: float) -> dict:
        tick = np.asarray(self.tick_latency) * 1e3
        lat = self.job_latencies()
        waits = [j.t_admit - j.t_arrive for j in self.done + list(self.running.values()) if j.t_admit is not None]
        per_job_p99 = [float(np.percentile(j.latency_ms, 99)) for j in self.done if j.latency_ms is not None
                       and len(j.latency_ms)]
        pct = lambda a, q: round(float(np.percentile(a, q)), 4) if len(a) else None
        return {
            "ticks": len(tick), "graph_steps": self.graph_steps,
            "step_latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "samples": int(len(lat)),
                                "source": "device step-end stamps, per job-step, gaps between steps included"},
            "step_latency_ms_median_job_p99": round(float(np.median(per_job_p99)), 4) if per_job_p99 else None,
            "tick_latency_ms_host": {"p50": pct(tick, 50), "p99": pct(tick, 99)},
            "atom_steps_per_s": round(self.atom_steps / wall, 1) if wall > 0 else 0.0,
            "gpu_busy_fraction": round(self.busy / wall, 4) if wall > 0 else 0.0,
            "jobs_done": len(self.done), "jobs_running": len(self.running), "jobs_queued": len(self.queue),
            "mean_admission_wait_s": round(float(np.mean(waits)), 3) if waits else None,
            "faulted_jobs": [j.jid for j in self.failed],
            "faults": [{"job": j.jid, "atoms": j.n_atoms, "at_step": j.done_steps, "error": j.error}
                       for j in self.failed],
        }
