"""Checkpoint / restore on the GPU (prism_md_checkpoint_*): a trajectory
interrupted, saved and restored -- into a fresh engine, into the same engine
after it ran on, or from a batch slot into a single-system engine -- continues
bitwise as the uninterrupted one, across pair-list rebuilds and re-sorts;
checkpoints of another system or other parameters are refused."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system(pm):
    top = pm.Topology.generate(3000, seed=17)
    p = pm.default_params(seed=123, stream_id=4, resort_every=2)  # re-sorts often
    return top, p


def test_restore_continues_bitwise(pm, system):
    top, p = system
    with pm.FusedEngine(top, p) as a:
        a.step(230)
        blob = a.checkpoint()
        st0 = a.stats()
        a.step(170)
        xa, va, sa = a.get_positions(), a.get_velocities(), a.stats()
        assert sa["pair_rebuilds"] > st0["pair_rebuilds"] + 2  # rebuilds (and re-sorts) after the save
        # a fresh engine from the checkpoint
        with pm.FusedEngine(top, p) as b:
            b.restore(blob)
            assert b.stats()["steps_done"] == 230
            b.step(170)
            assert np.array_equal(b.get_positions(), xa) and np.array_equal(b.get_velocities(), va)
            assert b.stats()["pair_rebuilds"] == sa["pair_rebuilds"]
        # the same engine, rewound
        a.restore(blob)
        a.step(170)
        assert np.array_equal(a.get_positions(), xa) and np.array_equal(a.get_velocities(), va)


def test_batch_slot_to_single_engine(pm, system):
    top, p = system
    other = pm.Topology.generate(1200, seed=18)
    with pm.Ensemble([other, top], [pm.default_params(seed=5), p]) as ens:
        ens.step(120, 10)
        blob = ens.checkpoint(1)
        ens.step(80, 10)
        xe = ens.get_positions(1)
    with pm.FusedEngine(top, p) as eng:
        eng.restore(blob)
        eng.step(80)
        assert np.array_equal(eng.get_positions(), xe)


def test_foreign_checkpoints_are_refused(pm, system):
    top, p = system
    with pm.FusedEngine(top, p) as a:
        a.step(20)
        blob = a.checkpoint()
    other = pm.Topology.generate(3000, seed=19)
    with pm.FusedEngine(other, p) as b:
        with pytest.raises(pm.PrismMdError) as ex:
            b.restore(blob)
        assert ex.value.rc == pm.E_INVALID and "different topology" in str(ex.value)
    with pm.FusedEngine(top, pm.default_params(seed=124, stream_id=4, resort_every=2)) as c:
        with pytest.raises(pm.PrismMdError) as ex:
            c.restore(blob)
        assert "different run parameters" in str(ex.value)
        with pytest.raises(pm.PrismMdError) as ex:
            c.restore(blob[:-8])
        assert ex.value.rc == pm.E_INVALID
        bad = bytearray(blob)
        bad[0] ^= 0xFF
        with pytest.raises(pm.PrismMdError):
            c.restore(bytes(bad))
        c.step(5)  # the engine is still usable after refused restores


def test_step_force_equals_recomputed_force(pm, system):
    """The step adds the nonbonded partials and bonded slots straight into
    the fixed-point accumulators; a separate evaluation (nonbonded, bonded,
    gather) at the same positions must give the same integers, so replacing
    the carried force by a recomputed one mid-run leaves the trajectory
    bitwise unchanged -- also on the first step after a restore."""
    top, p = system
    with pm.FusedEngine(top, p) as a, pm.FusedEngine(top, p) as b:
        a.step(90)
        b.step(90)
        fa = a.compute_forces_fixed()  # overwrites a's carried force with the gather's
        assert np.any(fa != 0)
        for e in (a, b):
            e.step(75)
        assert np.array_equal(a.get_positions(), b.get_positions())
        assert np.array_equal(a.get_velocities(), b.get_velocities())
        # energies use the energy variant of the same evaluation
        a.energies()
        for e in (a, b):
            e.step(40)
        assert np.array_equal(a.get_positions(), b.get_positions())
