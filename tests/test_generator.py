"""The seeded synthetic-chain generator against the workload description:
SAW heavy chain with 1.53 A bonds and ~111 degree angles in a sphere at
0.1 atoms/A^3, 1-2 H per heavy atom at 1.09 A, HMR H = 3.024 amu with the
donor reduced to match (input.rs:2482-2565), parameter ranges, neutral
charge, exclusions 1-2/1-3/1-4 and exact 1-4 exceptions scaled 1/2 (LJ) and
1/1.2 (Coulomb)."""
from __future__ import annotations

import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def chain(pm):
    return pm.Topology.generate(3000, seed=5)


def test_sizes_and_determinism(pm, chain):
    assert chain.n_atoms == 3000
    again = pm.Topology.generate(3000, seed=5)
    for k in ("positions", "charges", "bond_atoms", "dihedral_k", "exception_sigma"):
        assert np.array_equal(chain[k], again[k])
    other = pm.Topology.generate(3000, seed=6)
    assert not np.array_equal(chain["positions"], other["positions"])


def test_geometry(chain):
    x = chain["positions"].reshape(-1, 3).astype(np.float64)
    b = chain["bond_atoms"].reshape(-1, 2)
    r = np.linalg.norm(x[b[:, 0]] - x[b[:, 1]], axis=1)
    m = chain["masses"]
    heavy = m > 4.0
    hh = heavy[b[:, 0]] & heavy[b[:, 1]]
    assert np.all(chain["bond_r0"][hh] == np.float32(1.53)) and np.all(chain["bond_r0"][~hh] == np.float32(1.09))
    assert np.abs(r[hh] - 1.53).max() < 0.1  # built at 1.53, then relaxed against clashes
    assert np.abs(r[~hh] - 1.09).max() < 1e-5  # X-H snapped to the constraint length
    # sphere at 0.1 atoms / A^3
    R = (3 * chain.n_atoms / (4 * math.pi * 0.1)) ** (1 / 3)
    assert chain.gen_stats["sphere_radius_A"] == pytest.approx(R, rel=1e-6)
    dist = np.linalg.norm(x, axis=1)  # sphere centred at the origin
    assert dist[heavy].max() < R + 1.0  # heavy atoms confined (relaxation may push slightly out)
    assert dist.max() < R + 2.5
    # heavy-heavy-heavy equilibrium angles around 111 degrees (taken from the built geometry)
    a = chain["angle_atoms"].reshape(-1, 3)
    t0 = np.degrees(chain["angle_theta0"])
    hhh = heavy[a[:, 0]] & heavy[a[:, 1]] & heavy[a[:, 2]]
    assert abs(np.median(t0[hhh]) - 111.0) < 3.0


def test_hydrogens_and_hmr(chain):
    m = chain["masses"]
    heavy = m > 4.0
    assert np.allclose(m[~heavy], 3.024, atol=1e-4)
    nh = chain["cluster_n_hydrogens"]
    assert nh.min() >= 1 and nh.max() <= 3
    assert (~heavy).sum() == nh.sum()
    # every heavy atom carries 1-2 hydrogens
    assert len(nh) == heavy.sum() and nh.max() <= 2
    # donor (C, N or O) mass reduced by n_H * (3.024 - 1.008)
    donor = m[chain["cluster_central"]] + nh * (3.024 - 1.008)
    assert np.all(np.min(np.abs(donor[:, None] - np.array([12.011, 14.007, 15.999])[None]), axis=1) < 1e-3)
    # constraint lengths are the X-H bond lengths
    d = chain["cluster_bond_lengths"].reshape(-1, 3)
    assert np.allclose(d[d > 0], 1.09, atol=1e-6)


def test_parameter_ranges_and_charge(chain):
    q, s, e = chain["charges"], chain["sigma"], chain["epsilon"]
    assert abs(float(q.astype(np.float64).sum())) < 1e-3
    assert np.abs(q).max() <= 0.8 + 0.02  # U(-0.8, 0.8) draws shifted by the neutralising mean
    assert s.min() >= 1.0 - 1e-6 and s.max() <= 3.5 + 1e-6
    assert e.min() >= 0.01 - 1e-7 and e.max() <= 0.2 + 1e-7
    k = chain["bond_k"]
    assert k.min() >= 300 and k.max() <= 500
    ka = chain["angle_k"]
    assert ka.min() >= 40 and ka.max() <= 80
    n = chain["dihedral_periodicity"]
    assert n.min() >= 1 and n.max() <= 4
    ph = chain["dihedral_phase"]
    assert np.all(np.isclose(ph, 0.0) | np.isclose(ph, math.pi, atol=1e-6))


def test_exclusions_and_exceptions(chain):
    n = chain.n_atoms
    off, idx = chain["exclusion_offsets"], chain["exclusion_atoms"]
    sets = [set(idx[off[i]:off[i + 1]].tolist()) for i in range(n)]
    assert all(i in sets[j] for i in range(n) for j in sets[i])  # symmetric
    b = chain["bond_atoms"].reshape(-1, 2)
    assert all(int(j) in sets[int(i)] for i, j in b)
    ex = chain["exception_atoms"].reshape(-1, 2)
    assert all(int(j) in sets[int(i)] for i, j in ex)
    q, s, e = chain["charges"].astype(np.float64), chain["sigma"], chain["epsilon"]
    i, j = ex[:, 0], ex[:, 1]
    assert np.allclose(chain["exception_charge_product"], q[i] * q[j] / 1.2, rtol=1e-5, atol=1e-7)
    assert np.allclose(chain["exception_sigma"], 0.5 * (s[i] + s[j]), rtol=1e-6)
    assert np.allclose(chain["exception_epsilon"], 0.5 * np.sqrt(e[i] * e[j]), rtol=1e-5)


def test_no_clashes(chain):
    st = chain.gen_stats
    assert st["min_nonbonded_distance_A"] > 1.0
