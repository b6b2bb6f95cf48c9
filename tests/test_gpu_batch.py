"""Batched ensemble on the GPU: systems of uneven size in one engine, stepped
through a captured hipGraph, each bitwise identical to the same system run
alone (positions, velocities, fixed-point forces), with independent
activation."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [(700, 21), (2600, 22), (1500, 23), (5200, 24)]


@pytest.fixture(scope="module")
def systems(pm):
    tops, starts = [], []
    for n, seed in SIZES:
        top = pm.Topology.generate(n, seed=seed)
        with pm.FusedEngine(top) as eng:
            eng.minimize(800, 1.0)
            starts.append(eng.get_positions())
        tops.append(top)
    params = [pm.default_params(seed=100 + k, stream_id=k, temperature_K=290.0 + 10 * k,
                                end_temperature_K=290.0 + 10 * k) for k in range(len(tops))]
    return tops, starts, params


def _alone(pm, top, x, p, steps):
    with pm.FusedEngine(top, p) as eng:
        eng.set_state(x=x)
        eng.step(steps)
        return eng.get_positions(), eng.get_velocities(), eng.stats()


def test_batch_matches_alone_bitwise(pm, systems):
    tops, starts, params = systems
    with pm.Ensemble(tops, params) as ens:
        for s, x in enumerate(starts):
            ens.set_state(s, x=x)
        ens.step(250, graph_steps=10)  # 25 graph replays
        ens.step(7, graph_steps=0)     # and eager steps
        for s, top in enumerate(tops):
            xa, va, sa = _alone(pm, top, starts[s], params[s], 257)
            xb, vb = ens.get_positions(s), ens.get_velocities(s)
            sb = ens.stats(s)
            print(f"[system {s}] n={top.n_atoms} rebuilds alone {sa['pair_rebuilds']} batch {sb['pair_rebuilds']}",
                  flush=True)
            assert sb["steps_done"] == 257 and sb["status"] == 0
            assert np.array_equal(xa, xb) and np.array_equal(va, vb), s


def test_inactive_system_is_frozen_and_others_unaffected(pm, systems):
    tops, starts, params = systems
    with pm.Ensemble(tops, params) as ens:
        for s, x in enumerate(starts):
            ens.set_state(s, x=x)
        ens.step(40, graph_steps=10)
        x1 = ens.get_positions(1)
        ens.set_active(1, False)
        ens.step(60, graph_steps=10)
        assert np.array_equal(ens.get_positions(1), x1)
        assert ens.stats(1)["steps_done"] == 40
        ens.set_active(1, True)
        ens.step(30, graph_steps=10)
        assert ens.stats(1)["steps_done"] == 70 and ens.stats(0)["steps_done"] == 130
        # the untouched system followed its own trajectory throughout
        xa, va, _ = _alone(pm, tops[3], starts[3], params[3], 130)
        assert np.array_equal(xa, ens.get_positions(3))
        assert all(ens.stats(s)["status"] == 0 for s in range(len(tops)))


def test_batch_forces_match_oracle(pm, oracle, systems):
    tops, starts, params = systems
    with pm.Ensemble(tops, params) as ens:
        for s, x in enumerate(starts):
            ens.set_state(s, x=x)
        for s, top in enumerate(tops):
            F, E = ens.compute_forces(s)
            Fo, Eo = oracle.forces(top.with_positions(starts[s]), starts[s].astype(np.float64), 31, True)
            r = np.sqrt(((F - Fo) ** 2).sum() / (Fo ** 2).sum())
            assert r <= 1e-5, (s, r)


def test_large_arena_contribution_layout_matches_alone(pm, systems):
    """A batch of 21 x 5,200 atoms (an arena above 100,000 atoms) stores the
    bonded contributions term-major, a system alone atom-major; the forces
    are the same integers and the trajectories stay bitwise equal."""
    tops, starts, _ = systems
    top, x0 = tops[3], starts[3]
    n = 21
    params = [pm.default_params(seed=300 + k, stream_id=k) for k in range(n)]
    with pm.Ensemble([top] * n, params) as ens:
        for s in range(n):
            ens.set_state(s, x=x0)
        fb = ens.compute_forces(5)[0]
        ens.step(60, graph_steps=10)
        for s in (0, 5, 20):
            with pm.FusedEngine(top, params[s]) as eng:
                eng.set_state(x=x0)
                if s == 5:
                    fa, _ = eng.compute_forces()
                    assert np.array_equal(fa, fb)
                eng.step(60)
                assert np.array_equal(eng.get_positions(), ens.get_positions(s)), s
                assert np.array_equal(eng.get_velocities(), ens.get_velocities(s)), s
