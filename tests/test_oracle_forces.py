"""FP64 oracle force terms, pinned without the oracle's own formulas:

- central finite differences of the oracle's own energies (the reference's
  pattern, lambda_fragment.rs:14790-14860) on a generated system, per term;
- invariants: sum of forces = 0 and sum of torques = 0 per term;
- textbook special cases: 2-atom bond, angle at and off theta0, torsion at
  phi = 0 / pi / pi/2, LJ minimum at 2^(1/6) sigma, Coulomb pair, potential
  shift V(rc) = 0, exceptions without cutoff;
- a regression fixture (tests/golden/oracle_small.npz).
The FD check is shown to reject a planted sign error.
"""
from __future__ import annotations

import math
import os

import numpy as np
import pytest

from conftest import GOLDEN
from toy import make_topology

TERMS = {1: "bond", 2: "angle", 4: "dihedral", 8: "nonbonded", 16: "exception"}
COULOMB = 332.0636


def energy(oracle, top, x, mask):
    return oracle.forces(top, x, mask, False)[1][list(TERMS).index(mask)]


def fd_check(oracle, top, x, mask, F, idx, h=1e-5):
    errs = []
    for k in idx:
        xp = x.copy().reshape(-1)
        xm = x.copy().reshape(-1)
        xp[k] += h
        xm[k] -= h
        fd = -(energy(oracle, top, xp, mask) - energy(oracle, top, xm, mask)) / (2 * h)
        errs.append(abs(fd - F.reshape(-1)[k]) / max(1.0, abs(fd)))
    return max(errs)


@pytest.mark.parametrize("mask", list(TERMS))
def test_finite_differences(oracle, small_top, mask):
    x = small_top["positions"].reshape(-1, 3).astype(np.float64)
    F, E = oracle.forces(small_top, x, mask, False)
    idx = np.random.default_rng(mask).choice(x.size, 40, replace=False)
    assert fd_check(oracle, small_top, x, mask, F, idx) < 2e-6
    if mask == 1:  # a planted sign error must fail the same check
        assert fd_check(oracle, small_top, x, mask, -F, idx[:5]) > 0.5


@pytest.mark.parametrize("mask", list(TERMS))
def test_force_and_torque_sums(oracle, small_top, mask):
    x = small_top["positions"].reshape(-1, 3).astype(np.float64)
    F, _ = oracle.forces(small_top, x, mask, True)
    scale = np.abs(F).sum() + 1e-30
    assert np.abs(F.sum(axis=0)).max() < 1e-10 * scale
    c = x - x.mean(axis=0)
    assert np.abs(np.cross(c, F).sum(axis=0)).max() < 1e-9 * scale * (np.abs(c).max() + 1)


def test_two_atom_bond(pm, oracle):
    t = make_topology(pm, [[0, 0, 0], [1.7, 0, 0]], bonds=[(0, 1, 1.53, 400.0)], exclusions=[[1], [0]])
    F, E = oracle.forces(t, t["positions"].reshape(-1, 3), 1, False)
    d = 1.7 - 1.53
    assert E[0] == pytest.approx(400.0 * d * d, rel=1e-6)  # U = k (r - r0)^2, no 1/2
    assert F[1, 0] == pytest.approx(-2 * 400.0 * d, rel=1e-5)
    assert F[0, 0] == pytest.approx(2 * 400.0 * d, rel=1e-5)


def test_angle_at_and_off_theta0(pm, oracle):
    th0 = float(np.float32(math.radians(111.0)))  # the topology stores float
    for dth in (0.0, 0.1):
        th = th0 + dth
        pos = [[1.5, 0, 0], [0, 0, 0], [1.5 * math.cos(th), 1.5 * math.sin(th), 0]]
        t = make_topology(pm, pos, angles=[(0, 1, 2, th0, 60.0)], exclusions=[[1, 2], [0, 2], [0, 1]])
        F, E = oracle.forces(t, np.array(pos, np.float64), 2, False)
        assert E[1] == pytest.approx(60.0 * dth * dth, abs=1e-9)
        if dth == 0.0:
            assert np.abs(F).max() < 1e-5
        else:  # torque on atom 2 about the vertex: |F| * r = dU/dtheta = 2 k dth
            assert np.linalg.norm(F[2]) * 1.5 == pytest.approx(2 * 60.0 * dth, rel=1e-4)


@pytest.mark.parametrize("phi,energy_k", [(0.0, 2.0), (math.pi, 0.0), (math.pi / 2, 1.0)])
def test_torsion_closed_form(pm, oracle, phi, energy_k):
    # atoms i-j-k-l with j-k along x; l rotated by phi about the x axis
    pos = [[0, 1, 0], [0, 0, 0], [1.5, 0, 0], [1.5, math.cos(phi), math.sin(phi)]]
    t = make_topology(pm, pos, dihedrals=[(0, 1, 2, 3, 1, 0.0, 1.25)])
    F, E = oracle.forces(t, np.array(pos, np.float64), 4, False)
    assert E[2] == pytest.approx(1.25 * energy_k, abs=1e-9)  # V = k (1 + cos phi)
    if energy_k != 1.0:  # extrema: no force
        assert np.abs(F).max() < 1e-6
    else:  # dV/dphi = -k at pi/2: lever arm of atom l about the j-k axis is 1
        assert np.linalg.norm(F[3]) == pytest.approx(1.25, rel=1e-6)


def test_lj_minimum_and_shift(pm, oracle):
    sig, eps = 3.0, 0.25  # exact in float
    rmin = 2 ** (1 / 6) * sig
    t = make_topology(pm, [[0, 0, 0], [rmin, 0, 0]], sigma=[sig, sig], epsilon=[eps, eps])
    F, E = oracle.forces(t, np.array([[0, 0, 0], [rmin, 0, 0]], np.float64), 8, False)
    assert np.abs(F).max() < 1e-9
    shift = 4 * eps * ((sig / 10) ** 12 - (sig / 10) ** 6)
    assert E[3] == pytest.approx(-eps - shift, rel=1e-9)
    # V(rc) = 0 after the shift; beyond rc nothing
    for r, want in [(10.0 - 1e-9, 0.0), (10.5, 0.0)]:
        F, E = oracle.forces(t, np.array([[0, 0, 0], [r, 0, 0]], np.float64), 8, False)
        assert abs(E[3] - want) < 1e-9


def test_coulomb_pair(pm, oracle):
    t = make_topology(pm, [[0, 0, 0], [4.0, 0, 0]], charges=[0.5, -0.7], sigma=[0.0, 0.0])
    F, E = oracle.forces(t, np.array([[0, 0, 0], [4.0, 0, 0]], np.float64), 8, False)
    qq = 0.5 * -0.7
    assert E[3] == pytest.approx(COULOMB * qq * (1 / 4.0 - 1 / 10.0), rel=1e-6)
    assert F[1, 0] == pytest.approx(COULOMB * qq / 16.0, rel=1e-6)  # attraction pulls atom 1 to -x


def test_exception_has_no_cutoff(pm, oracle):
    r, qq, sig, eps = 11.0, 0.3, 3.2, 0.1
    t = make_topology(pm, [[0, 0, 0], [r, 0, 0]], exclusions=[[1], [0]], exceptions=[(0, 1, qq, sig, eps)])
    F, E = oracle.forces(t, np.array([[0, 0, 0], [r, 0, 0]], np.float64), 31, False)
    want = 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6) + COULOMB * qq / r
    assert E[4] == pytest.approx(want, rel=1e-6)
    assert E[3] == 0.0  # excluded from the nonbonded term


def test_excluded_pair_skipped(pm, oracle):
    t = make_topology(pm, [[0, 0, 0], [3.0, 0, 0]], charges=[1, 1], exclusions=[[1], [0]])
    F, E = oracle.forces(t, np.array([[0, 0, 0], [3.0, 0, 0]], np.float64), 8, False)
    assert E[3] == 0.0 and np.abs(F).max() == 0.0


def test_regression_fixture(oracle, small_top):
    g = np.load(os.path.join(GOLDEN, "oracle_small.npz"))
    assert np.array_equal(g["positions"], small_top["positions"]), "generator output changed"
    x = small_top["positions"].reshape(-1, 3).astype(np.float64)
    for mask in TERMS:
        F, E = oracle.forces(small_top, x, mask, True)
        assert np.allclose(F, g[f"F_{mask}"], rtol=1e-12, atol=1e-12)
        assert np.allclose(E, g[f"E_{mask}"], rtol=1e-12, atol=1e-12)
