"""The autotuner's parameters change only the tile layout, never the
physics: forces with any sort-cell size, re-sort interval and tiles per
nonbonded unit agree with the defaults to FP32 rounding and with the FP64
oracle; the tuner runs on the device and returns a measured choice for each
bucket."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_layout_parameters_do_not_change_forces(pm, oracle):
    top = pm.Topology.generate(5000, seed=23)
    x = top["positions"].reshape(-1, 3)
    Fo, _ = oracle.forces(top, x.astype(np.float64))
    base = None
    for cell, resort, chunk in [(2.0, 8, 0), (1.5, 4, 8), (3.0, 16, 32), (2.0, 8, 13)]:
        p = pm.default_params(sort_cell_A=cell, resort_every=resort, chunk_tiles=chunk)
        with pm.FusedEngine(top, p) as eng:
            F, E = eng.compute_forces()
            assert np.array_equal(eng.pairs(1), oracle.pairs(top, x, x, 1))
            eng.step(30)  # the step's direct accumulation with this layout
            assert eng.stats()["status"] == 0
        r = float(np.sqrt(((F - Fo) ** 2).sum() / (Fo ** 2).sum()))
        print(f"[layout] cell {cell} resort {resort} chunk {chunk}: rel RMS vs oracle {r:.2e}", flush=True)
        assert r <= 1e-5
        if base is None:
            base = F
        else:
            assert np.sqrt(((F - base) ** 2).sum() / (base ** 2).sum()) <= 2e-6


def test_tuner_on_device(pm):
    from prism_md import autotune
    table = autotune.tune(pm, [(800, 1200)], steps=60, graph_steps=10, copies=2,
                          sort_cells=(1.5, 3.0), resort_every=(4, 16), chunk_tiles=(8, 24))
    e = table[0]
    assert e["sort_cell_A"] in (1.5, 3.0) and e["resort_every"] in (4, 16) and e["chunk_tiles"] in (8, 24)
    assert len(e["trials"]) == 6 and all(t["us_per_step"] > 0 for t in e["trials"])
    assert e["us_per_step"] == min(t["us_per_step"] for t in e["trials"][4:])
