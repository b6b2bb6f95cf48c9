"""A system read from PrismPrepTopology JSON runs on the GPU exactly as the
in-memory topology it was written from: same pair list, same fixed-point
forces, bitwise the same trajectory."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_json_topology_runs_bitwise_like_generated(pm, tmp_path):
    top = pm.Topology.generate(2000, seed=13)
    path = tmp_path / "chain.json"
    top.to_json(str(path), hmr_applied=True)
    loaded = pm.Topology.from_json(str(path))
    out = []
    for t in (top, loaded):
        with pm.FusedEngine(t, pm.default_params(seed=99)) as eng:
            pairs = eng.pairs(1)
            F = eng.compute_forces_fixed()
            eng.step(60)
            out.append((pairs, F, eng.get_positions(), eng.get_velocities()))
    (p0, f0, x0, v0), (p1, f1, x1, v1) = out
    assert np.array_equal(p0, p1)
    assert np.array_equal(f0, f1)
    assert np.array_equal(x0, x1) and np.array_equal(v0, v1)
