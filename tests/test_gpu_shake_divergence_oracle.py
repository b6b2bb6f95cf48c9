"""A latched SHAKE fault of a generated system is the trajectory's, not the GPU's.

The 25,000-atom chain at 4 fs without a thermostat heats and, after a few
thousand steps, reaches a configuration whose constraints SHAKE cannot
restore.  Handed the GPU's own positions and velocities 40 steps before that
step, the FP64 oracle (Gauss-Seidel SHAKE, FP64 forces) fails within a step
of the same step: the fault belongs to the dynamics of the workload, which
FP32 forces and the device's Newton SHAKE follow faithfully.
"""
import re

import numpy as np
import pytest

GEN = 0x50524953 + 1
LEAD = 40


def _nve(pm):
    top = pm.Topology.generate(25000, seed=GEN)
    p = pm.default_params(dt_ps=0.004, gamma_ps=0.0, equilibration_steps=0, com_remove_every=0)
    return top, p


def _start(pm, eng, top):
    eng.minimize(2000, 1.0)
    eng.set_state(v=pm.maxwell_boltzmann(top["masses"], 300.0, 3))
    eng.step(1)  # RATTLE projects the drawn velocities


@pytest.mark.gpu
def test_nve_shake_divergence_step_matches_oracle(pm, oracle):
    top, p = _nve(pm)
    fault_step = None
    with pm.FusedEngine(top, p) as eng:
        _start(pm, eng, top)
        for _ in range(200):
            try:
                eng.step(100)
            except pm.PrismMdError as ex:
                m = re.search(r"SHAKE, atom \d+, step (\d+)", str(ex))
                assert m, str(ex)
                fault_step = int(m.group(1))
                break
    assert fault_step is not None, "no SHAKE fault within 20,000 steps"
    print(f"[fault] GPU SHAKE fault at step {fault_step}", flush=True)
    assert fault_step > LEAD
    with pm.FusedEngine(top, p) as eng:  # the same trajectory, stopped LEAD steps before
        _start(pm, eng, top)
        eng.step(fault_step - LEAD - 1)
        assert eng.stats()["steps_done"] == fault_step - LEAD
        x = eng.get_positions().astype(np.float64)
        v = eng.get_velocities().astype(np.float64)
    t = top.with_positions(x)
    failed_at = None
    for k in range(LEAD + 3):
        step = fault_step - LEAD + k
        try:
            x, v, _ = oracle.run(t, p, x, v, step, 1)
        except RuntimeError:
            failed_at = step
            break
    print(f"[fault] FP64 oracle from step {fault_step - LEAD}: fails at {failed_at}", flush=True)
    assert failed_at is not None and abs(failed_at - fault_step) <= 1, (failed_at, fault_step)
