"""Writes the committed fixtures of tests/golden/.

    python tests/golden/make_golden.py

philox_kat.json
    The Random123 known-answer vectors for philox4x32 with 10 rounds, as the
    reference quotes them (scripts/physics_refs/rng_streams.py:82-88).
director_table.json
    Temperature and timestep-hold table of the reference's director test
    (crates/prism-nhs/tests/protocol_director_thermal_contract.cpp:13-66): start
    83 K, end 317 K, cold hold to step 10, ramp to 30, warm hold to 60, then a
    ramp-down of 0 or 20 steps and a warm tail or cold return, for steps
    1..1007.  The rows are the test's expected() values; `hold` marks the steps
    where the test expects the base dt (adaptive dt halves it elsewhere).
oracle_small.npz
    Regression vectors of the FP64 oracle on a 192-atom generated system
    (per-term forces and energies, both pair-list modes).  They pin nothing on
    their own; the oracle itself is pinned by tests/test_oracle_*.py.  They
    catch an unintended change of the oracle or of the generator.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
M32 = 0xFFFFFFFF

PHILOX_KAT = [  # counter, key, expected output
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((M32, M32, M32, M32), (M32, M32), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
]


def director_rows(ramp_down_steps: int, warm_tail: bool):
    rows = []
    for step in range(1, 1008):
        if step < 10:
            t = 83.0
        elif step < 30:
            t = 83.0 + (step - 10) / 20.0 * 234.0
        elif step < 60:
            t = 317.0
        elif step < 60 + ramp_down_steps:
            t = 317.0 - (step - 60) / float(ramp_down_steps) * 234.0
        else:
            t = 317.0 if warm_tail else 83.0
        hold = step < 10 or (30 <= step < 60) or step >= 60 + ramp_down_steps
        rows.append([step, t, int(hold)])
    return rows


def write_director():
    cases = []
    for ramp_down, warm in ((0, True), (20, False), (0, False)):
        cases.append({"start_K": 83.0, "end_K": 317.0, "cold_hold_end": 10, "ramp_end": 30,
                      "warm_hold_end": 60, "ramp_down_end": 60 + ramp_down, "warm_tail": int(warm),
                      "rows": director_rows(ramp_down, warm)})
    with open(os.path.join(HERE, "director_table.json"), "w") as f:
        json.dump({"source": "crates/prism-nhs/tests/protocol_director_thermal_contract.cpp:13-66",
                   "tolerance_K": 1e-4, "cases": cases}, f)


def write_philox():
    with open(os.path.join(HERE, "philox_kat.json"), "w") as f:
        json.dump({"source": "Random123 kat_vectors (philox4x32, 10 rounds), "
                             "scripts/physics_refs/rng_streams.py:82-88",
                   "vectors": [{"ctr": list(c), "key": list(k), "out": list(o)} for c, k, o in PHILOX_KAT]}, f,
                  indent=1)


def write_oracle_small():
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "prism-q_amd"))
    import prism_md as pm
    from oracle import pyoracle as O
    t = pm.Topology.generate(192, seed=11)
    x = t["positions"].reshape(-1, 3).astype(np.float64)
    out = {"positions": t["positions"]}
    for k, mask in enumerate((1, 2, 4, 8, 16)):
        F, E = O.forces(t, x, mask, True)
        out[f"F_{mask}"] = F
        out[f"E_{mask}"] = E
    xf = t["positions"].reshape(-1, 3)
    out["pairs_0"] = O.pairs(t, xf, xf, 0)
    out["pairs_1"] = O.pairs(t, xf, xf, 1)
    np.savez_compressed(os.path.join(HERE, "oracle_small.npz"), **out)


if __name__ == "__main__":
    write_philox()
    write_director()
    write_oracle_small()
    print("fixtures written to", HERE)
