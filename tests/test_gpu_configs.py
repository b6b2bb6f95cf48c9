"""GPU parity at the benchmark's own configurations (the sizes the smaller
parity tests do not reach), against the FP64 oracle and against each
system run alone:

- the 250,000-atom globule (generator seed 0x50524953 + 2, 12 A list):
  brute-force pair list bit-exact and per-term forces within 1e-5 relative
  RMS, after minimisation and after 200 steps -- the size at which the
  term-major contribution slots, 16-tile work units and 512-tile rows are in
  use;
- the 256-system ensemble of the benchmark (sizes log-uniform 500-20,000,
  1.33 M atoms, per-system parameters varied as an autotuner would per size
  bucket, term-major arena, thousands of integrator workgroups so the
  step-end handshake runs its two levels): eight systems spanning the sizes
  against the oracle (forces) and against their runs alone (bitwise), with
  COM removal every 10 steps (the handshake's publish path) and one system
  forced into a constraint fault that halts it alone;
- a serving pool with the benchmark's slot classes (1,024 / 4,096 / 16,384 /
  51,200 atoms) holding jobs of 350 to 48,000 atoms: a slot's forces and one
  step of its own Philox noise against the oracle, and the largest job
  bitwise equal to its run alone.
Each test with its fixture stays within about a minute and a half.
"""
from __future__ import annotations

import concurrent.futures as cf
import os

import numpy as np
import pytest
from stepcheck import compare_step

pytestmark = pytest.mark.gpu

SEED = 0x50524953
TERMS = [(1, "bond"), (2, "angle"), (4, "dihedral"), (8, "nonbonded"), (16, "exception"), (31, "all")]


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / max(1e-300, (b ** 2).sum())))


def _workers():
    return max(1, min(16, os.cpu_count() or 1))


def _check_forces(F_of, oracle, top, x, tag):
    """Per-term relative RMS <= 1e-5 against the FP64 oracle on the same
    FP32 positions (cutoff membership decided in FP32 on both sides)."""
    t = top.with_positions(x)
    x64 = x.astype(np.float64)
    for mask, name in TERMS:
        F = F_of(mask)
        Fo, _ = oracle.forces(t, x64, mask, True)
        r = rel_rms(F.astype(np.float64), Fo)
        print(f"[{tag}] {name:9s} rel RMS {r:.2e}", flush=True)
        assert r <= 1e-5, (tag, name, r)


# ---------------------------------------------------------------- 250k globule
@pytest.fixture(scope="module")
def globule(pm):
    top = pm.Topology.generate(250000, seed=SEED + 2)
    eng = pm.FusedEngine(top, pm.default_params())
    eng.minimize(1000, 1.0)
    print(f"[250k] created and minimised: {eng.stats()}", flush=True)
    yield top, eng
    eng.close()


@pytest.mark.timeout(600)
def test_250k_pairs_after_minimisation(globule, oracle):
    top, eng = globule
    x = eng.get_positions()
    got = eng.pairs(1)
    want = oracle.pairs(top, x, x, 1)
    print(f"[250k] {len(got)} pairs (GPU) / {len(want)} (oracle)", flush=True)
    assert got.shape == want.shape and np.array_equal(got, want)


@pytest.mark.timeout(600)
def test_250k_forces_after_minimisation(globule, oracle):
    top, eng = globule
    x = eng.get_positions()
    _check_forces(lambda m: eng.compute_forces(m)[0], oracle, top, x, "250k min")


@pytest.mark.timeout(600)
def test_250k_after_200_steps(globule, oracle):
    top, eng = globule
    eng.step(200)
    st = eng.stats()
    assert st["status"] == 0 and st["steps_done"] == 200
    x = eng.get_positions()
    _check_forces(lambda m: eng.compute_forces(m)[0], oracle, top, x, "250k 200 steps")
    got = eng.pairs(1)
    want = oracle.pairs(top, x, x, 1)
    assert np.array_equal(got, want)


# ---------------------------------------------------------- 256-system batch
N_SYS = 256
FAULTY = 3       # forced into a constraint fault (one SHAKE sweep at 1e-14)
CHECK = None     # eight systems spanning the sizes, set by the fixture


@pytest.fixture(scope="module")
def ensemble(pm):
    rng = np.random.default_rng(SEED + 3)  # the benchmark's ensemble draw
    sizes = np.exp(rng.uniform(np.log(500), np.log(20000), N_SYS)).astype(int)
    temps = rng.uniform(280.0, 340.0, N_SYS)
    with cf.ThreadPoolExecutor(_workers()) as ex:  # the generator releases the interpreter lock
        tops = list(ex.map(lambda k: pm.Topology.generate(int(sizes[k]), seed=SEED + 4 + k), range(N_SYS)))
    # per-bucket layout parameters, as the autotuner assigns them
    bucket = np.digitize(sizes, [1400, 3900, 8000])
    cells = {0: 1.5, 1: 2.0, 2: 3.0, 3: 2.0}
    resort = {0: 4, 1: 8, 2: 16, 3: 8}
    params = []
    for k in range(N_SYS):
        kw = dict(seed=SEED + 1000 + k, stream_id=k, temperature_K=float(temps[k]),
                  end_temperature_K=float(temps[k]), com_remove_every=10, chunk_tiles=16,
                  sort_cell_A=cells[int(bucket[k])], resort_every=resort[int(bucket[k])])
        if k == FAULTY:
            kw.update(constraint_max_iter=1, shake_tol=1e-14, rattle_tol=1e-14)
        params.append(pm.default_params(**kw))
    order = np.argsort(sizes)
    check = [int(order[i]) for i in np.linspace(0, N_SYS - 1, 8).astype(int) if int(order[i]) != FAULTY]
    ens = pm.Ensemble(tops, params)
    print(f"[ensemble] {N_SYS} systems, {ens.n_atoms()} arena atoms", flush=True)
    yield tops, params, sizes, check, ens
    ens.close()


@pytest.mark.timeout(600)
def test_ensemble_forces_match_oracle(ensemble, oracle):
    tops, params, sizes, check, ens = ensemble
    for s in check:
        x = ens.get_positions(s)
        F, _ = ens.compute_forces(s)
        Fo, _ = oracle.forces(tops[s].with_positions(x), x.astype(np.float64), 31, True)
        r = rel_rms(F.astype(np.float64), Fo)
        print(f"[ensemble] system {s} ({sizes[s]} atoms): force rel RMS {r:.2e}", flush=True)
        assert r <= 1e-5, (s, r)


@pytest.mark.timeout(600)
def test_ensemble_matches_alone_with_a_faulted_member(pm, ensemble):
    """100 steps (10 graph replays) with COM removal every 10 steps: system
    FAULTY latches a SHAKE fault in its first step and halts there; every other
    system runs all 100 steps, and the sampled ones are bitwise equal to their
    runs alone."""
    tops, params, sizes, check, ens = ensemble
    with pytest.raises(pm.PrismMdError) as ex:
        ens.step(100, 10)
    assert ex.value.rc == pm.E_CONSTRAINT and f"system {FAULTY}" in str(ex.value)
    st = ens.stats(FAULTY)
    assert st["status"] != 0 and st["steps_done"] == 1
    ens.set_active(FAULTY, False)  # the fault no longer fails the batch's calls
    for s in check:
        assert ens.stats(s)["steps_done"] == 100 and ens.stats(s)["status"] == 0
        with pm.FusedEngine(tops[s], params[s]) as eng:
            eng.step(100)
            same = np.array_equal(eng.get_positions(), ens.get_positions(s)) and \
                np.array_equal(eng.get_velocities(), ens.get_velocities(s))
        print(f"[ensemble] system {s} ({sizes[s]} atoms) bitwise alone: {same}", flush=True)
        assert same, s
    ens.step(20, 10)  # the others run on
    assert ens.stats(check[0])["steps_done"] == 120


# ------------------------------------------------------------- serving pool
CLASSES = [1024, 4096, 16384, 51200]
JOBS = [(350, 61), (3000, 62), (12000, 63), (48000, 64)]


@pytest.fixture(scope="module")
def served(pm):
    template = pm.Topology.generate(4000, seed=SEED + 5)
    caps = [pm.slot_caps_for(template, atoms=c, scale=1.3) for c in CLASSES for _ in range(2)]
    pool = pm.ServingPool(caps)
    with cf.ThreadPoolExecutor(_workers()) as ex:
        tops = list(ex.map(lambda j: pm.Topology.generate(j[0], seed=SEED + j[1]), JOBS))
    jobs = []
    for k, top in enumerate(tops):
        p = pm.default_params(seed=SEED + 900 + k, stream_id=k, temperature_K=280.0 + 15 * k,
                              end_temperature_K=280.0 + 15 * k)
        slot = next(i for i in range(len(caps)) if not pool.occupied(i) and pool.fits(i, top))
        pool.admit(slot, top, p)
        jobs.append((slot, top, p))
    print(f"[pool] {len(caps)} slots, jobs in slots {[j[0] for j in jobs]}", flush=True)
    yield pool, jobs
    pool.close()


@pytest.mark.timeout(600)
def test_pool_slot_forces_and_step_match_oracle(oracle, served):
    pool, jobs = served
    pool.step(30, 10)
    for slot, top, p in jobs[1:3]:
        x = pool.get_positions(slot)
        _check_forces(lambda m: pool.compute_forces(slot, m)[0], oracle, top, x, f"pool slot {slot}")
    # one step of every job with the device's own noise; job 2 against the oracle
    slot, top, p = jobs[2]
    xs, vs = pool.get_positions(slot), pool.get_velocities(slot)
    step0 = pool.stats(slot)["steps_done"]
    pool.step(1, 0)
    xg, vg = pool.get_positions(slot), pool.get_velocities(slot)
    xo, vo, _ = oracle.run(top.with_positions(xs), p, xs, vs, step0, 1)
    compare_step(top, p, xg, vg, xo, vo, 2e-5, 1e-4, f"pool slot {slot} step {step0}")


@pytest.mark.timeout(600)
def test_pool_largest_job_bitwise_alone(pm, served):
    pool, jobs = served
    slot, top, p = jobs[3]
    n = pool.stats(slot)["steps_done"]
    with pm.FusedEngine(top, p) as eng:
        eng.step(n)
        assert np.array_equal(eng.get_positions(), pool.get_positions(slot))
        assert np.array_equal(eng.get_velocities(), pool.get_velocities(slot))
