"""GPU parity through the C ABI against the FP64 oracle on identical inputs.

Bars (the tolerances are written in each test):
- pair lists: bit-exact in both modes (reference cell scheme, brute force);
- forces: per-term relative RMS <= 1e-5 vs the FP64 oracle, energies per
  term <= 1e-5 relative (nonbonded cutoff membership decided in FP32 on both
  sides, so the evaluated pair set is the same);
- fixed-point forces and trajectories: bitwise reproducible run to run;
- integrator: one step from the GPU's own state with the Langevin normals
  fed in explicitly, positions within 2e-5 A and velocities within 1e-4
  relative RMS of the oracle's FP64 step.
Sizes: the 2,000-atom chain and the 25,000-atom chain of the benchmark.
"""
from __future__ import annotations

import numpy as np
import pytest
from stepcheck import compare_step

pytestmark = pytest.mark.gpu

TERMS = [(1, "bond"), (2, "angle"), (4, "dihedral"), (8, "nonbonded"), (16, "exception")]


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / max(1e-300, (b ** 2).sum())))


@pytest.fixture(scope="module", params=[2000, 25000], ids=["2k", "25k"])
def system(request, pm):
    n = request.param
    top = pm.Topology.generate(n, seed=7)
    eng = pm.FusedEngine(top, pm.default_params())
    print(f"[system] n={n} created", flush=True)
    yield top, eng
    eng.close()


@pytest.mark.parametrize("mode", [0, 1], ids=["reference", "brute"])
def test_pairs_bit_exact_at_creation(system, oracle, mode):
    top, eng = system
    x = top["positions"].reshape(-1, 3)
    got = eng.pairs(mode)
    want = oracle.pairs(top, x, x, mode)
    assert got.shape == want.shape and np.array_equal(got, want)


def _check_terms(eng, oracle, top_x, x):
    for mask, name in TERMS + [(31, "all")]:
        F, E = eng.compute_forces(mask)
        Fo, Eo = oracle.forces(top_x, x.astype(np.float64), mask, True)
        r = rel_rms(F.astype(np.float64), Fo)
        print(f"[forces] {name:9s} rel RMS {r:.2e}", flush=True)
        assert r <= 1e-5, (name, r)
        if mask != 31:
            k = [m for m, _ in TERMS].index(mask)
            assert abs(E[k] - Eo[k]) <= 1e-5 * max(1.0, abs(Eo[k])), (name, E[k], Eo[k])
        else:
            assert abs(E.sum() - Eo.sum()) <= 1e-5 * abs(Eo.sum())


def test_forces_minimized(system, oracle):
    top, eng = system
    e, fmax = eng.minimize(2000, 1.0)
    assert np.isfinite(e) and fmax < 50.0
    x = eng.get_positions()
    _check_terms(eng, oracle, top.with_positions(x), x)
    g = eng.pairs(1)
    assert np.array_equal(g, oracle.pairs(top, x, x, 1))


def test_forces_thermal(system, oracle):
    top, eng = system
    eng.step(200)
    x = eng.get_positions()
    _check_terms(eng, oracle, top.with_positions(x), x)
    assert eng.stats()["status"] == 0


def test_newton_third_law(system, pm):
    """Each pair is evaluated from both of its atoms with the separation
    negated, which gives exactly opposite pair forces, so the nonbonded forces
    sum to zero up to the FP32 rounding of the partial sums (relative to the
    largest atom force)."""
    top, eng = system
    F, _ = eng.compute_forces(pm.TERM_NONBONDED)
    net = np.abs(F.astype(np.float64).sum(axis=0)).max()
    print(f"[newton] |sum F| {net:.3e}  max|F| {np.abs(F).max():.3e}", flush=True)
    assert net <= 1e-5 * np.abs(F).max() * np.sqrt(len(F))


def test_fixed_point_forces_reproducible(system):
    top, eng = system
    a = eng.compute_forces_fixed()
    b = eng.compute_forces_fixed()
    assert np.array_equal(a, b)


def test_integrator_explicit_noise_per_step(system, oracle):
    top, eng = system
    n = top.n_atoms
    nz = np.random.default_rng(1).standard_normal((3, 2, n, 3)).astype(np.float32)
    step0 = eng.stats()["steps_done"]
    xs, vs = eng.get_positions(), eng.get_velocities()
    for k in range(3):
        eng.step(1, noise=nz[k:k + 1])
        xg, vg = eng.get_positions(), eng.get_velocities()
        xo, vo, _ = oracle.run(top.with_positions(xs), eng.params, xs, vs, step0 + k, 1, nz[k:k + 1])
        compare_step(top, eng.params, xg, vg, xo, vo, 2e-5, 1e-4, f"step {k}")
        xs, vs = xg, vg


def test_run_to_run_bitwise(pm):
    top = pm.Topology.generate(2000, seed=9)
    out = []
    for _ in range(2):
        with pm.FusedEngine(top, pm.default_params(seed=1234)) as eng:
            eng.minimize(500, 1.0)
            eng.step(300)
            out.append((eng.get_positions(), eng.get_velocities(), eng.compute_forces_fixed()))
    for a, b in zip(*out):
        assert np.array_equal(a, b)


def test_step_seam_equals_separate_halves(pm):
    """A multi-step call joins each step's second half and the next step's
    first half in one launch (the step seam); single-step calls run them as
    separate kernels.  300 steps cross COM removals (every 100 steps, the
    seam's publishing path), rebuilds and a re-sort: bitwise the same state,
    work ledger and rebuild count either way."""
    top = pm.Topology.generate(2000, seed=9)
    out = []
    for one_at_a_time in (False, True):
        with pm.FusedEngine(top, pm.default_params(seed=77)) as eng:
            eng.minimize(500, 1.0)
            if one_at_a_time:
                for _ in range(300):
                    eng.step(1)
            else:
                eng.step(300)
            st = eng.stats()
            assert st["status"] == 0 and st["steps_done"] == 300, st
            out.append((eng.get_positions(), eng.get_velocities(), eng.energies()["work"], st["pair_rebuilds"]))
    a, b = out
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert a[2] == b[2] and a[3] == b[3], (a[2:], b[2:])


def test_seed_changes_trajectory(pm):
    top = pm.Topology.generate(2000, seed=9)
    res = []
    for seed in (1, 2):
        with pm.FusedEngine(top, pm.default_params(seed=seed)) as eng:
            eng.step(20)
            res.append(eng.get_positions())
    assert not np.array_equal(res[0], res[1])
