"""The oracle's Philox4x32-10 stream, Box-Muller normals and the director
schedule, pinned by the reference's own known-answer data.

- Philox: Random123 KATs (scripts/physics_refs/rng_streams.py:82-88) and the
  counter packing of keyed_draw (rng_streams.py:75-79), checked through an
  independent pure-Python Philox that is itself pinned by the same KATs.
- Box-Muller: the published transform applied to the KAT-pinned integers;
  the sample moments of a large draw.
- Director: the temperature table and dt hold flags of
  crates/prism-nhs/tests/protocol_director_thermal_contract.cpp:13-66 and the
  equilibration friction ramp of fused_engine.rs:15370-15378.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

M32 = 0xFFFFFFFF


def philox_py(ctr, key):
    """Philox4x32-10 from the Random123 paper (Salmon et al., SC'11)."""
    c0, c1, c2, c3 = ctr
    k0, k1 = key
    for r in range(10):
        if r:
            k0 = (k0 + 0x9E3779B9) & M32
            k1 = (k1 + 0xBB67AE85) & M32
        p0 = 0xD2511F53 * c0
        p1 = 0xCD9E8D57 * c2
        c0, c1, c2, c3 = (p1 >> 32) ^ c1 ^ k0, p1 & M32, (p0 >> 32) ^ c3 ^ k1, p0 & M32
    return (c0, c1, c2, c3)


@pytest.fixture(scope="module")
def kat():
    with open(os.path.join(GOLDEN, "philox_kat.json")) as f:
        return json.load(f)["vectors"]


def test_philox_python_matches_kat(kat):
    for v in kat:
        assert philox_py(v["ctr"], v["key"]) == tuple(v["out"])


def test_oracle_philox_matches_kat(oracle, kat):
    for v in kat:
        assert oracle.philox(v["ctr"], v["key"]) == tuple(v["out"])


def test_langevin_keying_and_box_muller(oracle):
    # counter = (atom, step lo, (step hi16 << 16) | purpose << 8 | epoch 0, 0),
    # key = (seed lo, stream): rng_streams.py:75-79
    for seed, stream, atom, step, purpose in [(0x50524953, 0, 0, 0, 1), (7, 3, 12345, 2**33 + 17, 2),
                                              (2**32 - 1, 2**32 - 1, 2**31, 2**47, 3)]:
        ctr = (atom, step & M32, ((step >> 32) << 16) | (purpose << 8), 0)
        r = philox_py(ctr, (seed & M32, stream))
        s = 2.0 ** -32
        a = math.sqrt(-2.0 * math.log((r[0] + 1) * s))
        b = math.sqrt(-2.0 * math.log((r[2] + 1) * s))
        want = [a * math.cos(2 * math.pi * r[1] * s), a * math.sin(2 * math.pi * r[1] * s),
                b * math.cos(2 * math.pi * r[3] * s)]
        got = oracle.langevin_normals(seed, stream, atom, step, purpose)
        assert np.allclose(got, want, rtol=0, atol=1e-12)


def test_langevin_normal_moments(oracle):
    z = np.array([oracle.langevin_normals(99, 1, a, 5, 1) for a in range(20000)]).ravel()
    assert abs(z.mean()) < 0.02
    assert abs(z.var() - 1.0) < 0.03
    # distinct purposes / steps give independent draws
    z2 = np.array([oracle.langevin_normals(99, 1, a, 5, 2) for a in range(2000)]).ravel()
    assert abs(np.corrcoef(z[:6000], z2)[0, 1]) < 0.05


@pytest.fixture(scope="module")
def director():
    with open(os.path.join(GOLDEN, "director_table.json")) as f:
        return json.load(f)


def test_director_temperature_table(pm, oracle, director):
    tol = director["tolerance_K"]
    for case in director["cases"]:
        p = pm.default_params(temperature_K=case["start_K"], end_temperature_K=case["end_K"],
                              cold_hold_end=case["cold_hold_end"], ramp_end=case["ramp_end"],
                              warm_hold_end=case["warm_hold_end"], ramp_down_end=case["ramp_down_end"],
                              warm_tail=case["warm_tail"], equilibration_steps=0)
        for step, T, hold in case["rows"]:
            t, _, h = oracle.schedule(p, step)
            assert abs(t - T) < tol, (case["ramp_down_end"], step, t, T)
            assert h == hold, (step, h, hold)


def test_equilibration_friction_ramp(pm, oracle):
    p = pm.default_params(gamma_ps=10.0, equilibration_steps=10000, equilibration_gamma_ps=1000.0)
    for s, want in [(0, 1000.0), (10000, 10.0), (20000, 10.0)]:
        assert abs(oracle.schedule(p, s)[1] - want) < 1e-9
    # half-way: 1000 e^-1.5 + 10 (1 - e^-1.5)
    e = math.exp(-1.5)
    assert abs(oracle.schedule(p, 5000)[1] - (1000.0 * e + 10.0 * (1 - e))) < 1e-9
    # monotone decreasing over the ramp
    g = [oracle.schedule(p, s)[1] for s in range(0, 10001, 250)]
    assert all(a > b for a, b in zip(g, g[1:]))


def test_constant_schedule_defaults(pm, oracle):
    p = pm.default_params(temperature_K=310.0, end_temperature_K=310.0, equilibration_steps=0)
    for s in (0, 1, 99, 10**6):
        t, g, h = oracle.schedule(p, s)
        assert t == 310.0 and g == p.gamma_ps and h == 1
